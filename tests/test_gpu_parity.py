"""GPU parity: the HIP path (through the C ABI) against the reference's golden
vectors and the CPU oracle on the same inputs and the same sample stream.

Bar: integer and index results (entry tets, tets crossed, triangles, n_avg)
exactly equal; FP64 results within 1e-10 relative per value (the FP64 path
keeps the reference's operation order, so in practice they are bit-equal,
which is also reported).
"""
import hashlib
import sys

import numpy as np
import pytest

from sdtfe_amd import capi, workloads
from sdtfe_amd.tessellate import to_particle_data, triangulate

pytestmark = pytest.mark.gpu

RTOL_FP64 = 1e-10


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    d = np.abs(a - b)
    return np.max(d / np.maximum(np.abs(b), 1e-300)) if a.size else 0.0


def _contains(pd, tri, tet, x, y):
    """Is (x, y) inside (inclusive) a forward triangle of tet `tet`?"""
    for row in tri[tri[:, 6] == tet]:
        a, b, c = pd[row[0], :2], pd[row[1], :2], pd[row[2], :2]
        ok = True
        for p, q, o in ((a, b, c), (b, c, a), (c, a, b)):
            ex, ey = q[0] - p[0], q[1] - p[1]
            d_ins = (o[0] - p[0]) * ey - (o[1] - p[1]) * ex
            d_qry = (x - p[0]) * ey - (y - p[1]) * ex
            ok &= not (d_ins * d_qry < 0.0)
        if ok:
            return True
    return False


@pytest.fixture(scope="module", autouse=True)
def _device():
    n = capi.device_count()
    assert n >= 1, "no HIP device visible to libsdtfe_amd"


@pytest.mark.parametrize("case", ["tet4", "pts8", "pts64", "dyadic"])
def test_field_matches_reference(golden, case):
    _say(f"[gpu] field {case}")
    g = golden("g1_tiny.npz")
    pd = to_particle_data(g[f"{case}_pos"])
    with capi.Tessellation(pd, g[f"{case}_tetra"], p_mass=0.75) as t:
        np.testing.assert_array_equal(t.vertex_density(), g[f"{case}_rho"])
        np.testing.assert_array_equal(t.gradients(), g[f"{case}_nabla"])
        tri = t.triangles()
        ref = g[f"{case}_tri"]
        np.testing.assert_array_equal(tri, ref)


def test_entry_and_rays_match_reference(golden):
    _say("[gpu] rays g2 (fp64)")
    g = golden("g2_rays.npz")
    pd = to_particle_data(g["pos"])
    with capi.Tessellation(pd, g["tetra"]) as t:
        np.testing.assert_array_equal(t.vertex_density(), g["rho"])
        entry = t.locate(g["x"], g["y"])
        n_c = int(g["n_crafted"])
        # stream samples: the same entry tet as the reference's walk
        assert (entry[:-n_c] == g["entry"][:-n_c]).all()
        # crafted rays sit exactly on projected vertices/edges, where several
        # forward triangles contain the point; any containing one is valid
        diff = np.nonzero(entry != g["entry"])[0]
        tri = t.triangles()
        for i in diff:
            assert _contains(pd, tri, int(entry[i]), g["x"][i], g["y"][i]), i
        _say(f"  entry differs on {diff.size} crafted on-edge rays (all valid)")
        val, crossed, st = t.trace(g["x"], g["y"], g["pixel"], g["sample"], float(g["off_z"]),
                                   float(g["lz"]), float(g["dxy"]), int(g["seed"]),
                                   entry=g["entry"])
        np.testing.assert_array_equal(crossed, g["crossed"])
        assert _rel(val, g["value"]) <= RTOL_FP64
        n_exact = int((val == g["value"]).sum())
        _say(f"  rays bit-equal {n_exact}/{val.size}, perturb steps {st['perturb_steps']}")
        assert st["perturb_steps"] > 100  # crafted rays exercised the perturbation


MAPS = __import__("tests.golden.make_golden", fromlist=["MAPS"]).MAPS


@pytest.mark.parametrize("name,G,S,L,Lz,cen,ds,m,seed", [pytest.param(*m, id=m[0]) for m in MAPS])
def test_maps_match_reference(golden, name, G, S, L, Lz, cen, ds, m, seed):
    _say(f"[gpu] map {name}")
    g = golden("g3_maps.npz")
    pd = to_particle_data(g["pos"])
    with capi.Tessellation(pd, g["tetra"], p_mass=m) as t:
        np.testing.assert_array_equal(t.vertex_density(), g[f"{name}_vertex_rho"])
        rho, st = t.map(G, S, box_len=L, box_len_z=Lz, center=cen, delta_sample=ds, seed=seed)
    assert _rel(rho, g[f"{name}_rho"]) <= RTOL_FP64
    _say(f"  bit-equal pixels {(rho == g[f'{name}_rho']).sum()}/{rho.size}")
    assert st["samples"] == G * G * S


def test_uniform_32768_matches_reference(golden):
    _say("[gpu] uniform-32768 256x256x4 vs reference golden")
    g = golden("g4_c1.npz")
    pos = workloads.uniform(32768, 0)
    tetra = triangulate(pos)
    if hashlib.sha256(tetra.tobytes()).hexdigest() != str(g["tetra_sha"]):
        pytest.skip("scipy/Qhull tessellates differently from the fixture")
    spec = workloads.C1["spec"]
    with capi.Tessellation(to_particle_data(pos), tetra, p_mass=spec.p_mass) as t:
        rho, st, navg = t.map(256, 4, want_navg=True, box_len=spec.box_len,
                              box_len_z=spec.box_len_z, center=spec.center,
                              delta_sample=spec.delta_sample, seed=int(g["seed"]))
    assert _rel(rho, g["rho"]) <= RTOL_FP64
    _say(f"  bit-equal pixels {(rho == g['rho']).sum()}/{rho.size}; steps {st['tet_steps']}")
    assert st["samples"] == 256 * 256 * 4 and st["failed"] == 0


def test_clustered_map_matches_oracle(oracle):
    """A clustered set (halos with steep density) at a size the oracle
    finishes in seconds; per-pixel values and n_avg against the oracle."""
    _say("[gpu] clustered-65536 256x256x8 vs oracle")
    pos = workloads.clustered(65536, seed=3)
    tetra = triangulate(pos)
    pd = to_particle_data(pos)
    G, S, seed = 256, 8, 21
    ref, st_o, navg_o = oracle.compute_density(pd.copy(), tetra, G, 1.0, 1.0, 1.0,
                                               (0.5, 0.5, 0.5), S, 0.0, seed, want_navg=True)
    with capi.Tessellation(pd, tetra) as t:
        rho, st, navg = t.map(G, S, want_navg=True, seed=seed)
    np.testing.assert_array_equal(navg, navg_o)
    assert _rel(rho, ref) <= RTOL_FP64
    assert st["tet_steps"] == st_o["tet_steps"]
    assert st["perturb_steps"] == st_o["perturb_steps"]


def test_dropin_compute_density(golden):
    """The reference entry point through the ABI: rho accumulates, column 3
    receives the vertex density, and the seed advances per call."""
    _say("[gpu] drop-in compute_density")
    g = golden("g3_maps.npz")
    name, G, S, L, Lz, cen, ds, m, seed = MAPS[0]
    pd = to_particle_data(g["pos"])
    rho = np.full(G * G, 0.5)
    capi.set_seed(seed)
    capi.compute_density(pd, g["tetra"], G, L, Lz, rho, m, cen, S, ds)
    np.testing.assert_array_equal(pd[:, 3], g[f"{name}_vertex_rho"])
    assert _rel(rho - 0.5, g[f"{name}_rho"]) <= 1e-9
    assert capi.lib().sdtfe_get_seed() == seed + 1
    # SDTFE_PRECISION=mixed selects the mixed walk for the drop-in
    import os
    pd = to_particle_data(g["pos"])
    mixed = np.zeros(G * G)
    capi.set_seed(seed)
    os.environ["SDTFE_PRECISION"] = "mixed"
    try:
        capi.compute_density(pd, g["tetra"], G, L, Lz, mixed, m, cen, S, ds)
    finally:
        os.environ.pop("SDTFE_PRECISION", None)
    with capi.Tessellation(to_particle_data(g["pos"]), g["tetra"], p_mass=m) as t:
        want, _ = t.map(G, S, box_len=L, box_len_z=Lz, center=cen, delta_sample=ds, seed=seed,
                        precision=capi.PREC_MIXED)
    np.testing.assert_array_equal(mixed, want)
    assert _rel(mixed, g[f"{name}_rho"]) <= 1e-5


@pytest.mark.parametrize("prec", [capi.PREC_FP64, capi.PREC_MIXED])
def test_map_device_accumulates_and_repeats(prec):
    _say(f"[gpu] map_device accumulate + repeatability, precision {prec}")
    pos = workloads.uniform(20000, 4)
    tetra = triangulate(pos)
    with capi.Tessellation(to_particle_data(pos), tetra) as t:
        host, _ = t.map(128, 4, seed=9, precision=prec)
        buf = capi.DeviceBuffer(128 * 128 * 8)
        buf.zero()
        p = t.params(128, 4, seed=9, precision=prec)
        t.map_device(p, buf)
        t.map_device(p, buf)
        dev = buf.to_host()
        buf.close()
    np.testing.assert_array_equal(dev, host + host)


def test_errors_are_reported():
    _say("[gpu] error paths")
    pos = workloads.uniform(200, 1)
    tetra = triangulate(pos)
    bad = tetra.copy()
    bad[3, 5] = tetra.shape[0] + 7
    with pytest.raises(capi.SdtfeError, match="out of range"):
        capi.Tessellation(to_particle_data(pos), bad)
    with capi.Tessellation(to_particle_data(pos), tetra) as t:
        with pytest.raises(capi.SdtfeError):
            t.map(0, 4)
        rho, st = t.map(16, 2, box_len=3.0)  # most samples outside the hull's shadow
        assert st["missed"] > 0 and np.isfinite(rho).all()


@pytest.mark.parametrize("S", [1, 5, 20, 300])
def test_sample_split_maps_match_oracle(oracle, S):
    """Maps whose samples are spread over 1, 4 or 16 lanes per pixel (by
    the sample count; 20 leaves the second round partly empty, 300 exceeds
    the split's window and runs one lane per pixel): FP64 against the oracle,
    mixed against FP64 with the same tets."""
    _say(f"[gpu] clustered-8192 48x48x{S}, lanes per pixel by sample count")
    pos = workloads.clustered(8192, seed=4)
    tetra = triangulate(pos)
    pd = to_particle_data(pos)
    G, seed = 48, 5
    ref, st_o, navg_o = oracle.compute_density(pd.copy(), tetra, G, 1.0, 1.0, 1.0,
                                               (0.5, 0.5, 0.5), S, 0.0, seed, want_navg=True)
    with capi.Tessellation(pd, tetra) as t:
        rho, st, navg = t.map(G, S, want_navg=True, seed=seed)
        r32, s32, n32 = t.map(G, S, want_navg=True, seed=seed, precision=capi.PREC_MIXED)
    np.testing.assert_array_equal(navg, navg_o)
    assert _rel(rho, ref) <= RTOL_FP64
    assert st["tet_steps"] == st_o["tet_steps"]
    assert st["samples"] == G * G * S
    np.testing.assert_array_equal(n32, navg)
    assert s32["tet_steps"] == st["tet_steps"]
    assert s32["samples"] == G * G * S
    assert _rel(r32, rho) <= 1e-5


@pytest.mark.parametrize("S,box_len_z,max_steps", [(1, 1.0, 0), (5, 1.0, 0), (20, 1.0, 0),
                                                   (1, 0.4, 0), (5, 0.4, 0), (20, 0.4, 0),
                                                   (5, 1.0, 40)])
def test_fp64_walks_agree(monkeypatch, S, box_len_z, max_steps):
    """The FP64 walk over face records (the default) and the walk over
    per-tet records (SDTFE_FP64_WALK=rec) give the same maps, n_avg and
    counters bit for bit, with 1, 4 and 16 lanes per pixel, over the whole
    hull, through a slab that cuts it, and with a step cap that stops some
    rays."""
    _say(f"[gpu] clustered-20000 64x64x{S} box_len_z {box_len_z} max_steps {max_steps}: "
         "face vs rec walk")
    pos = workloads.clustered(20000, seed=8)
    tetra = triangulate(pos)
    pd = to_particle_data(pos)
    out = {}
    with capi.Tessellation(pd, tetra) as t:
        for w in ("face", "rec"):
            monkeypatch.setenv("SDTFE_FP64_WALK", w)
            out[w] = t.map(64, S, want_navg=True, seed=3, box_len_z=box_len_z,
                           max_steps=max_steps)
    (a, sa, na), (b, sb, nb) = out["face"], out["rec"]
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(na, nb)
    counters = [k for k in sa if not k.startswith("t_")]
    assert {k: sa[k] for k in counters} == {k: sb[k] for k in counters}
    assert sa["tet_steps"] > 0
    if max_steps:
        assert sa["step_cap"] > 0


def test_lattice_map_matches_oracle(oracle):
    """The large-map layout (jittered lattice plus halos, slab of half the
    box height) at 32x32x16 particles: FP64 map, n_avg and tets against the
    oracle."""
    _say("[gpu] lattice-16384 96x96x8 vs oracle")
    pos = workloads.jittered_lattice(32, 32, 16, seed=6)
    tetra = triangulate(pos)
    pd = to_particle_data(pos)
    G, S, seed = 96, 8, 9
    ref, st_o, navg_o = oracle.compute_density(pd.copy(), tetra, G, 1.0, 0.5, 1.0,
                                               (0.5, 0.5, 0.25), S, 0.0, seed, want_navg=True)
    with capi.Tessellation(pd, tetra) as t:
        rho, st, navg = t.map(G, S, want_navg=True, seed=seed, box_len_z=0.5,
                              center=(0.5, 0.5, 0.25))
    np.testing.assert_array_equal(navg, navg_o)
    assert _rel(rho, ref) <= RTOL_FP64
    assert st["tet_steps"] == st_o["tet_steps"]


@pytest.mark.parametrize("G,S", [(1, 3), (7, 0), (7, 5), (50, 20)])
def test_small_and_ragged_maps(oracle, G, S):
    """Grids of 1 and 7 pixels and one not divisible by 4, and no samples
    at all.  The reference walks 4x4 tiles (src/dtfe.c:601-608), so it
    covers the first 4 floor(G / 4) rows and columns; there the FP64 map
    equals the oracle's, and everywhere the mixed map matches the FP64 one.
    With no samples every pixel receives 0 (src/dtfe.c:639-645)."""
    _say(f"[gpu] ragged map {G}x{G}x{S}")
    pos = workloads.clustered(4096, seed=8)
    tetra = triangulate(pos)
    pd = to_particle_data(pos)
    ref, _, navg_o = oracle.compute_density(pd.copy(), tetra, G, 1.0, 1.0, 1.0,
                                            (0.5, 0.5, 0.5), S, 0.0, 4, want_navg=True)
    with capi.Tessellation(pd, tetra) as t:
        rho, st, navg = t.map(G, S, want_navg=True, seed=4)
        r32, s32, n32 = t.map(G, S, want_navg=True, seed=4, precision=capi.PREC_MIXED)
    assert st["samples"] == G * G * S and s32["samples"] == G * G * S
    c = 4 * (G // 4)
    cov = np.zeros((G, G), dtype=bool)
    cov[:c, :c] = True
    cov = cov.ravel()
    np.testing.assert_array_equal(navg[cov], navg_o[cov])
    assert _rel(rho[cov], ref[cov]) <= RTOL_FP64
    np.testing.assert_array_equal(n32, navg)
    assert s32["tet_steps"] == st["tet_steps"]
    assert np.isfinite(rho).all() and _rel(r32, rho) <= 1e-5
    if S == 0:
        assert not rho.any() and not r32.any() and not navg.any()


@pytest.mark.parametrize("factor,rot", [(1.0, None), (0.6, [0.3, -0.7, 1.1])])
def test_pydtfe_surface_density(oracle, factor, rot):
    """The pydtfe mirror end to end (float32 positions, shuffle, mass /
    sample_factor, rotation, Qhull with the GPU tet assembly, then the
    drop-in compute_density) against the oracle's compute_density on the
    same pre-processed particles."""
    from sdtfe_amd import pydtfe
    _say(f"[gpu] pydtfe.compute_surface_density factor={factor} rot={rot is not None}")
    pos = workloads.clustered(30000, seed=17)
    cen, G, S = [0.5, 0.45, 0.5], 96, 6
    got = pydtfe.compute_surface_density(pos, cen, G, factor, 0.9, 0.7, 2.0, 0.02, S,
                                         rotation_angle=rot, seed=5)
    assert got.shape == (G, G)
    pd = pydtfe.convert_to_qhdata(np.asarray(pos, dtype=np.float32, order="F"))
    n_sh = int(np.float32(pos.shape[0]) * np.float32(factor))
    pydtfe.shuffle_buff(pd, pos.shape[0], n_sh, 5)
    if rot is not None:
        pydtfe.rotate3d(pd, n_sh, rot, cen)
    pd = np.ascontiguousarray(pd[:n_sh])
    mass = float(np.float32(2.0 / float(np.float32(factor))))
    want, _ = oracle.compute_density(pd, triangulate(pd), G, 0.9, 0.7, mass, cen, S, 0.02,
                                     5, n_threads=8)
    assert _rel(got.ravel(), want) <= RTOL_FP64
    np.testing.assert_array_equal(got.ravel(), want)
