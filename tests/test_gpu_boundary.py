"""The reference's C surface and Python package name through
libsdtfe_amd.so on the GPU: a C program written against the reference's
prototypes and linked against the library alone, the forward-hull and
gradient drop-ins against the reference's own outputs (G1, G3), failures of
void drop-ins reaching Python, and ``import pydtfe``."""
import os
import subprocess
import sys

import numpy as np
import pytest

from sdtfe_amd import capi
from sdtfe_amd.tessellate import to_particle_data

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MAPS = __import__("tests.golden.make_golden", fromlist=["MAPS"]).MAPS


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def build_caller(tmp_path):
    """gcc the C caller against libsdtfe_amd.so (no other object)."""
    libdir = os.path.dirname(capi.LIB_PATH)
    exe = str(tmp_path / "c_caller")
    subprocess.run(["gcc", "-O2", "-std=c99", "-Wall", os.path.join(ROOT, "tests", "c_caller.c"),
                    "-o", exe, "-L" + libdir, "-lsdtfe_amd", "-Wl,-rpath," + libdir],
                   check=True)
    return exe


def run_caller(exe, tmp_path, pd, tetra, G, L, Lz, m, cen, S, ds, seed):
    out = tmp_path / f"out_{G}_{seed}"
    out.mkdir()
    pd = np.ascontiguousarray(pd, dtype=np.float64)
    tetra = np.ascontiguousarray(tetra, dtype=np.int32)
    pd.tofile(out / "pos.bin")
    tetra.tofile(out / "tets.bin")
    args = [exe, str(out / "pos.bin"), str(pd.shape[0]), str(out / "tets.bin"),
            str(tetra.shape[0]), str(G), repr(L), repr(Lz), repr(m), repr(cen[0]), repr(cen[1]),
            repr(cen[2]), str(S), repr(ds), str(seed), str(out)]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    nt = tetra.shape[0]
    res = {
        "rho": np.fromfile(out / "rho.bin", dtype=np.float64),
        "vrho": np.fromfile(out / "vrho.bin", dtype=np.float64),
        "nabla": np.fromfile(out / "nabla.bin", dtype=np.float64).reshape(nt, 3),
        "tri": np.fromfile(out / "tri.bin", dtype=np.int32).reshape(-1, 8),
        "bitmap": np.fromfile(out / "bitmap.bin", dtype=np.uint32),
    }
    return res


def test_c_caller_map_matches_reference(golden, tmp_path):
    """compute_density called from C as the reference's CLI calls it: the map
    of G3 (written by the reference's own compute_density) within 1e-10 per
    pixel, the vertex densities bit for bit."""
    _say("[gpu] C caller: G3 map")
    exe = build_caller(tmp_path)
    g = golden("g3_maps.npz")
    name, G, S, L, Lz, cen, ds, m, seed = MAPS[0]
    res = run_caller(exe, tmp_path, to_particle_data(g["pos"]), g["tetra"], G, L, Lz, m, cen, S,
                     ds, seed)
    ref = g[f"{name}_rho"]
    rel = np.max(np.abs(res["rho"] - ref) / np.maximum(np.abs(ref), 1e-300))
    assert rel <= 1e-10, rel
    np.testing.assert_array_equal(res["vrho"], g[f"{name}_vertex_rho"])


@pytest.mark.parametrize("case", ["tet4", "pts8", "pts64", "dyadic"])
def test_c_caller_field_matches_reference(golden, tmp_path, case):
    """pre_compute_vol, gradients, find_forward_hull and copy_fh_triangles
    called from C give the reference's G1 outputs bit for bit."""
    _say(f"[gpu] C caller: G1 {case}")
    exe = build_caller(tmp_path)
    g = golden("g1_tiny.npz")
    pd = to_particle_data(g[f"{case}_pos"])
    res = run_caller(exe, tmp_path, pd, g[f"{case}_tetra"], 8, 1.0, 1.0, 0.75,
                     (0.5, 0.5, 0.5), 1, 0.0, 3)
    np.testing.assert_array_equal(res["vrho"], g[f"{case}_rho"])
    np.testing.assert_array_equal(res["nabla"], g[f"{case}_nabla"])
    np.testing.assert_array_equal(res["tri"], g[f"{case}_tri"])
    # bitmap: bit f of word t for each emitted face, nothing else
    bm = res["bitmap"]
    assert int(sum(bin(int(w)).count("1") for w in bm)) == res["tri"].shape[0]
    assert bm[-1] == 0


@pytest.mark.parametrize("case", ["pts64", "dyadic"])
def test_hull_and_gradient_dropins(golden, case):
    """The same drop-ins through ctypes; a bitmap already holding bits keeps
    them (set_bit semantics) and the count covers only new marks."""
    _say(f"[gpu] forward hull drop-ins {case}")
    g = golden("g1_tiny.npz")
    pd = to_particle_data(g[f"{case}_pos"])
    pd[:, 3] = g[f"{case}_rho"]
    tet = g[f"{case}_tetra"]
    n, bm = capi.find_forward_hull(tet, pd)
    ref = g[f"{case}_tri"]
    assert n == ref.shape[0]
    np.testing.assert_array_equal(capi.copy_fh_triangles(tet, pd, n, bm), ref)
    ptr = capi.lib().gradients(np.ascontiguousarray(tet), tet.shape[0], pd)
    capi.check_dropin("gradients")
    nab = capi._take_malloced(ptr, 3 * tet.shape[0], capi.C.c_double, np.float64)
    np.testing.assert_array_equal(nab.reshape(-1, 3), g[f"{case}_nabla"])


def test_bad_tetra_raises():
    """A void drop-in's failure reaches Python as SdtfeError (an index out of
    range is found on the GPU before any walk)."""
    _say("[gpu] bad tetra_data raises")
    pd = to_particle_data(np.random.default_rng(0).random((16, 3)))
    tet = np.zeros((2, 8), dtype=np.int32)
    tet[:, 4:] = -1
    tet[1, 2] = 99
    rho = np.zeros(64)
    with pytest.raises(capi.SdtfeError, match="out of range"):
        capi.compute_density(pd, tet, 8, 1.0, 1.0, rho, 1.0, (0.5, 0.5, 0.5), 1, 0.0)
    with pytest.raises(capi.SdtfeError, match="out of range"):
        capi.find_forward_hull(tet, pd)
    # a good call afterwards reports success again
    from sdtfe_amd import workloads
    from sdtfe_amd.tessellate import triangulate
    pos = workloads.uniform(256, 1)
    capi.compute_density(to_particle_data(pos), triangulate(pos), 8, 1.0, 1.0, rho, 1.0,
                         (0.5, 0.5, 0.5), 1, 0.0)
    assert capi.lib().sdtfe_dropin_status() == 0


def test_import_pydtfe(oracle):
    """`import pydtfe` as the reference's callers do
    (examples/add_particle_dtfe_density.py:27): the same functions, checked
    against the oracle on one small map."""
    _say("[gpu] import pydtfe")
    import pydtfe
    from sdtfe_amd import pydtfe as mirror
    from sdtfe_amd import workloads
    for f in ("compute_surface_density", "compute_3d_density", "interpolate_field2grid",
              "compute_particle_density"):
        assert getattr(pydtfe, f) is getattr(mirror, f)
    pos = workloads.clustered(3000, seed=2).astype(np.float32)
    rho = pydtfe.compute_surface_density(pos, [0.5, 0.5, 0.5], 32, 1.0, 1.0, 1.0, 1.0, 0.0, 2,
                                         seed=4)
    # the oracle on the same shuffled particles and sample stream
    pd = mirror.convert_to_qhdata(pos)
    mirror.shuffle_buff(pd, pd.shape[0], pd.shape[0], 4)
    from sdtfe_amd.tessellate import triangulate
    tetra = triangulate(pd)
    ref, _ = oracle.compute_density(pd, tetra, 32, 1.0, 1.0, 1.0, (0.5, 0.5, 0.5), 2, 0.0, 4)
    rel = np.max(np.abs(rho.ravel() - ref) / np.maximum(np.abs(ref), 1e-300))
    assert rel <= 1e-10, rel
