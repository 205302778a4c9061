"""The volumetric device code (sdtfe_amd/csrc/volume.hip) compiled for the
host and run on the reference's volumetric fixtures: the column form and the
ray + fold form (with and without bands, and with every column marked for
the column form to redo) must both give the reference's grid bit for bit.
Pins the kernels' logic without a GPU; the GPU tests check the device build.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from sdtfe_amd import build as sb
from sdtfe_amd import workloads
from sdtfe_amd.tessellate import to_particle_data, triangulate
from tests.test_walk_host import records

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def host_volume(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("hv") / "host_volume")
    cmd = [hipcc, "-x", "hip", "-std=c++17", "-O2", "-ffp-contract=off",
           f"--offload-arch={sb.ARCH}", "-I" + sb.CSRC, "-I" + os.path.join(sb.ROOT, "include"),
           "-w", os.path.join(HERE, "host_volume.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True)
    return exe


def run_volume(exe, tmp_path, pd, tetra, G, ss, L, cen, seed, entries, init, mark_all=0,
               band_rows=0):
    """Both forms of the host build; returns (column grid, ray grid,
    counters of each, marked columns)."""
    rho = pd[:, 3]
    nt = tetra.shape[0]
    nab = np.empty((nt, 3))
    from oracle import oracle as O
    nab[:] = O.gradients(tetra, pd)
    rec = records(pd[:, :3], tetra, rho, nab)
    inp = tmp_path / "vol.in"
    with open(inp, "wb") as f:
        f.write(np.array([nt, G, ss, mark_all, band_rows], dtype=np.int64).tobytes())
        f.write(np.array([L, *cen], dtype=np.float64).tobytes())
        f.write(np.array([seed], dtype=np.uint64).tobytes())
        f.write(rec.tobytes())
        f.write(np.ascontiguousarray(entries, dtype=np.int32).tobytes())
        f.write(np.ascontiguousarray(init, dtype=np.float64).tobytes())
    out = tmp_path / "vol.out"
    r = subprocess.run([exe, str(inp), str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    b = np.fromfile(out, dtype=np.uint8)
    cells = G ** 3
    grids = b[:16 * cells].view(np.float64)
    cnt = b[16 * cells:].view(np.int64)
    keys = ("samples", "missed", "failed", "tet_steps", "perturb_steps")
    return (grids[:cells], grids[cells:], dict(zip(keys, cnt[:5])), dict(zip(keys, cnt[5:10])),
            int(cnt[10]))


def _entries(oracle, pd, tetra, G, L, cen, ss):
    tri = oracle.triangles(tetra, pd)
    x, y = oracle.volume_ray_xy(G, L, cen, ss)
    return np.array([oracle.pt_loc_2d(tri, pd, float(a), float(b)) for a, b in zip(x, y)],
                    dtype=np.int32)


@pytest.mark.parametrize("mark_all,band_rows", [(0, 0), (0, 5), (1, 0)])
def test_device_volume_code_matches_reference(oracle, golden, host_volume, tmp_path,
                                              mark_all, band_rows):
    """G6's interpolate_to_3d_grid fixture (12^3, supersampling 3, a
    non-linear field into a non-zero grid, a box wider than the hull)."""
    import tests.golden.make_golden as M
    g = golden("g6_volume.npz")
    _, G, ss, L, cen, m, seed = next(v for v in M.VOLUMES if v[0] == "field")
    pd = to_particle_data(g["pos"])
    pd[:, 3] = M._field(g["pos"])
    tetra = g["tetra"]
    ent = _entries(oracle, pd, tetra, G, L, cen, ss)
    col, ray, cc, rc, marked = run_volume(host_volume, tmp_path, pd, tetra, G, ss, L, cen, seed,
                                          ent, g["field_init"], mark_all, band_rows)
    np.testing.assert_array_equal(col, g["field_grid"])
    np.testing.assert_array_equal(ray, g["field_grid"])
    assert cc == rc
    # every column with a ray inside the hull's shadow is marked (the others
    # have nothing to add)
    hit = (ent.reshape(G * G, ss * ss) >= 0).any(axis=1).sum()
    assert marked == (hit if mark_all else 0) and 0 < hit < G * G


@pytest.mark.parametrize("name,G,ss", [("crafted16", 16, 1), ("crafted8", 8, 2)])
def test_device_volume_code_perturbed_rays(golden, host_volume, tmp_path, name, G, ss):
    """Rays through vertices and along edges of a dyadic point set: both
    forms with the reference's own entry tets give its grid."""
    g = golden("g6_volume.npz")
    pd = to_particle_data(g["crafted_pos"])
    pd[:, 3] = g["crafted_rho"]
    col, ray, cc, rc, _ = run_volume(host_volume, tmp_path, pd, g["crafted_tetra"], G, ss, 1.0,
                                     (0.5, 0.5, 0.5), 7, g[f"{name}_entries"], np.zeros(G ** 3))
    np.testing.assert_array_equal(col, g[f"{name}_grid"])
    np.testing.assert_array_equal(ray, g[f"{name}_grid"])
    assert cc["perturb_steps"] == int(g[f"{name}_rand_calls"]) and cc == rc


def test_device_volume_code_clustered(oracle, host_volume, tmp_path):
    """A clustered set with halos, supersampling 2, against the oracle."""
    pos = workloads.clustered(20000, seed=31)
    tetra = triangulate(pos)
    pd = to_particle_data(pos)
    oracle.density(pd, tetra, 0.5)
    G, ss, L, cen = 20, 2, 0.8, (0.45, 0.5, 0.55)
    ent = _entries(oracle, pd, tetra, G, L, cen, ss)
    want, st = oracle.volume_columns(pd, tetra, oracle.gradients(tetra, pd),
                                     oracle.triangles(tetra, pd), G, L, cen, ss, seed=3,
                                     entries=ent, n_threads=4)
    col, ray, cc, rc, _ = run_volume(host_volume, tmp_path, pd, tetra, G, ss, L, cen, 3, ent,
                                     np.zeros(G ** 3), band_rows=7)
    np.testing.assert_array_equal(col, want)
    np.testing.assert_array_equal(ray, want)
    for k in cc:
        assert cc[k] == rc[k] == st[k], k
