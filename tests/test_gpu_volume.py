"""GPU parity of the volumetric path (interpolate_to_3d_grid and
compute_3d_density, src/dtfe.c:376-516, 657-735) through the C ABI.

Bar: bit-equal grids.  One lane owns one grid column and adds its rays'
terms in the reference's order, in FP64 with the reference's operations, so
the grid equals the reference's (golden fixtures written by the reference)
and the oracle's bit for bit, as do the counters.
"""
import os
import sys

import numpy as np
import pytest

from sdtfe_amd import capi, pydtfe, workloads
from sdtfe_amd.tessellate import to_particle_data, triangulate

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _device():
    assert capi.device_count() >= 1, "no HIP device visible to libsdtfe_amd"


@pytest.mark.parametrize("name", ["dens", "field"])
def test_volume_matches_reference(golden, name):
    import tests.golden.make_golden as M
    _say(f"[gpu] volume {name} vs reference fixture")
    g = golden("g6_volume.npz")
    _, G, ss, L, cen, m, seed = next(v for v in M.VOLUMES if v[0] == name)
    pd = to_particle_data(g["pos"])
    if m is None:
        pd[:, 3] = M._field(g["pos"])
        t = capi.Tessellation(pd, g["tetra"], given_field=True)
    else:
        t = capi.Tessellation(pd, g["tetra"], p_mass=m)
    with t:
        if m is not None:
            np.testing.assert_array_equal(t.vertex_density(), g[f"{name}_vertex_rho"])
        grid, st = t.volume(G, ss, grid=g[f"{name}_init"].copy(), box_len=L, center=cen,
                            seed=seed)
    np.testing.assert_array_equal(grid, g[f"{name}_grid"])
    assert st["samples"] == G * G * ss * ss and st["perturb_steps"] == 0


@pytest.mark.parametrize("name,G,ss", [("crafted16", 16, 1), ("crafted8", 8, 2)])
def test_volume_perturbed_rays_match_oracle(oracle, golden, name, G, ss):
    """Rays through vertices and along edges (the perturbation path): the
    GPU grid against the oracle's, which the reference pins on the same
    rays.  On a projected edge the GPU's entry lookup may pick the other
    triangle, so the oracle walks from the GPU's entries."""
    _say(f"[gpu] volume {name} (perturbed rays) vs oracle")
    g = golden("g6_volume.npz")
    pd = to_particle_data(g["crafted_pos"])
    pd[:, 3] = g["crafted_rho"]
    tetra = g["crafted_tetra"]
    x, y = oracle.volume_ray_xy(G, 1.0, (0.5, 0.5, 0.5), ss)
    with capi.Tessellation(pd, tetra, given_field=True) as t:
        entries = t.locate(x, y)
        grid, st = t.volume(G, ss, seed=7)
    want, st_o = oracle.volume_columns(pd, tetra, oracle.gradients(tetra, pd),
                                       oracle.triangles(tetra, pd), G, 1.0, (0.5, 0.5, 0.5),
                                       ss, seed=7, entries=entries, n_threads=4)
    _say(f"  perturbation steps {st['perturb_steps']}")
    assert st["perturb_steps"] == st_o["perturb_steps"] > 100
    assert st["tet_steps"] == st_o["tet_steps"]
    np.testing.assert_array_equal(grid, want)


@pytest.mark.parametrize("G,ss,L,cen", [(64, 2, 1.0, (0.5, 0.5, 0.5)),
                                        (45, 3, 0.6, (0.4, 0.55, 0.45)),
                                        (1, 2, 1.0, (0.5, 0.5, 0.5)),
                                        (96, 1, 1.1, (0.5, 0.5, 0.5))])
def test_clustered_volume_matches_oracle(oracle, G, ss, L, cen):
    """131,072 clustered particles (halos): density grid against the
    oracle's compute_3d_density, a ragged grid (45 is no multiple of the
    16 x 16 block), a one-cell grid and supersampling 1 (the column kernel,
    pydtfe's default) on a box wider than the hull included."""
    _say(f"[gpu] clustered volume G={G} ss={ss}")
    pos = workloads.clustered(131072, seed=12)
    tetra = triangulate(pos)
    pd = to_particle_data(pos)
    want, st_o = oracle.volume(pd.copy(), tetra, G, L, cen, ss, seed=4, p_mass=0.25,
                               n_threads=8)
    with capi.Tessellation(pd, tetra, p_mass=0.25) as t:
        grid, st = t.volume(G, ss, box_len=L, center=cen, seed=4)
    for k in ("samples", "missed", "failed", "tet_steps", "perturb_steps"):
        assert st[k] == st_o[k], k
    np.testing.assert_array_equal(grid, want)


def test_volume_device_accumulates():
    _say("[gpu] volume_device accumulate + repeatability")
    pos = workloads.uniform(20000, 4)
    tetra = triangulate(pos)
    with capi.Tessellation(to_particle_data(pos), tetra) as t:
        host, _ = t.volume(32, 2, seed=2)
        buf = capi.DeviceBuffer(32 ** 3 * 8)
        buf.zero()
        p = t.volume_params(32, 2, seed=2)
        t.volume_device(p, buf)
        once = buf.to_host().copy()
        t.volume_device(p, buf)
        twice = buf.to_host()
        buf.close()
    np.testing.assert_array_equal(once, host)
    # the second pass adds each term onto the first pass's values
    assert np.all(twice[host > 0] > host[host > 0])


def test_dropin_3d_entry_points(oracle):
    """compute_3d_density and interpolate_to_3d_grid with the reference's
    prototypes: column 3 receives the density, grids accumulate, the seed
    advances per call."""
    _say("[gpu] drop-in compute_3d_density / interpolate_to_3d_grid")
    rng = np.random.default_rng(5)
    pos = rng.random((3000, 3))
    tetra = triangulate(pos)
    pd = to_particle_data(pos)
    pd_o = pd.copy()
    G, ss = 20, 2
    start = rng.random(G ** 3)
    rho = start.copy()
    capi.set_seed(11)
    capi.compute_3d_density(pd, tetra, G, 1.0, rho, 0.5, (0.5, 0.5, 0.5), ss)
    assert capi.lib().sdtfe_get_seed() == 12
    want, _ = oracle.volume(pd_o, tetra, G, 1.0, (0.5, 0.5, 0.5), ss, seed=11, p_mass=0.5,
                            grid=start.copy(), n_threads=8)
    np.testing.assert_array_equal(pd[:, 3], pd_o[:, 3])
    np.testing.assert_array_equal(rho, want)
    field = to_particle_data(pos)
    field[:, 3] = np.cos(3 * pos[:, 0]) + pos[:, 2]
    grid = np.zeros(G ** 3)
    capi.interpolate_to_3d_grid(field, tetra, G, 1.1, grid, (0.5, 0.5, 0.5), 1)
    want, _ = oracle.volume(field.copy(), tetra, G, 1.1, (0.5, 0.5, 0.5), 1, seed=12,
                            n_threads=8)
    np.testing.assert_array_equal(grid, want)


def test_pydtfe_3d_functions(oracle):
    """The pydtfe mirror's compute_3d_density, interpolate_field2grid and
    compute_particle_density against the oracle on the same pre-processed
    particles (float32 positions, shuffle, mass / sample_factor)."""
    _say("[gpu] pydtfe 3D functions")
    rng = np.random.default_rng(9)
    pos = rng.random((4000, 3))
    cen, G = [0.5, 0.5, 0.5], 16
    got = pydtfe.compute_3d_density(pos, cen, G, 0.5, 1.0, 2.0, 2, seed=3)
    assert got.shape == (G, G, G)
    pd = pydtfe.convert_to_qhdata(np.asarray(pos, dtype=np.float32, order="F"))
    pydtfe.shuffle_buff(pd, 4000, 2000, 3)
    pd = np.ascontiguousarray(pd[:2000])
    want, _ = oracle.volume(pd, triangulate(pd), G, 1.0, cen, 2, seed=3,
                            p_mass=float(np.float32(2.0 / np.float32(0.5))), n_threads=8)
    np.testing.assert_array_equal(got.ravel(), want)
    val = np.sin(5 * pos[:, 1]).astype(np.float32)
    got = pydtfe.interpolate_field2grid(pos, val, cen, G, 1.0, 1.2, 1, seed=4)
    pd = pydtfe.convert_to_qhdata(np.asarray(pos, dtype=np.float32, order="F"))
    tet = triangulate(pd)
    pd[:, 3] = val
    want, _ = oracle.volume(pd, tet, G, 1.2, cen, 1, seed=4, n_threads=8)
    np.testing.assert_array_equal(got.ravel(), want)
    dens = pydtfe.compute_particle_density(pos, 3.0)
    pd = pydtfe.convert_to_qhdata(np.asarray(pos, dtype=np.float32, order="F"))
    want = oracle.density(pd, tet, 3.0).astype(np.float32)
    assert dens.dtype == np.float32
    np.testing.assert_array_equal(dens, want)
    with pytest.raises(ValueError, match="different length"):
        pydtfe.interpolate_field2grid(pos, val[:-1], cen, G, 1.0, 1.0)
    with pytest.raises(ValueError, match="not supported"):
        pydtfe.interpolate_field2grid(pos, val, cen, G, 0.5, 1.0)


@pytest.mark.parametrize("G,ss", [(48, 2), (21, 3)])
def test_volume_forms_agree(oracle, G, ss):
    """With supersampling > 1 the grid comes from the ray kernels (a lane per
    ray writes its terms, a lane per cell adds them in order); the column
    kernel (a lane per column walks its rays in order), the redo path
    (every column marked and recomputed by the column kernel) and bands of
    5 grid rows give the same grid and counters bit for bit, and so does
    the oracle."""
    _say(f"[gpu] volume forms G={G} ss={ss}")
    pos = workloads.clustered(65536, seed=21)
    tetra = triangulate(pos)
    pd = to_particle_data(pos)
    want, st_o = oracle.volume(pd.copy(), tetra, G, 0.9, (0.5, 0.48, 0.52), ss, seed=6,
                               p_mass=1.0, n_threads=8)
    start = np.random.default_rng(2).random(G ** 3)
    res = {}
    with capi.Tessellation(pd, tetra) as t:
        for form in ("", "columns", "redo", "bands"):
            os.environ["SDTFE_VOLUME_FORM"] = form
            os.environ["SDTFE_VOLUME_BAND_ROWS"] = "5" if form == "bands" else ""
            try:
                res[form] = t.volume(G, ss, grid=start.copy(), box_len=0.9,
                                     center=(0.5, 0.48, 0.52), seed=6)
            finally:
                os.environ.pop("SDTFE_VOLUME_FORM", None)
                os.environ.pop("SDTFE_VOLUME_BAND_ROWS", None)
    want_acc, _ = oracle.volume(pd.copy(), tetra, G, 0.9, (0.5, 0.48, 0.52), ss, seed=6,
                                p_mass=1.0, grid=start.copy(), n_threads=8)
    for form, (grid, st) in res.items():
        np.testing.assert_array_equal(grid, want_acc, err_msg=form)
        for k in ("samples", "missed", "failed", "tet_steps", "perturb_steps"):
            assert st[k] == st_o[k], (form, k)
    assert np.isfinite(want).all()
