"""The map at the workloads' stated sizes, each against the oracle:

* the clustered map: 1,048,576 particles, 2048^2 pixels x 16 samples, FP64
  against the oracle on sampled 4-row bands, mixed against FP64 everywhere;
* the large single map: 8,388,608 particles (256 x 256 x 128 jittered
  lattice plus 5% halos), 56.5M tets (so the face records take 64-bit
  offsets), 4096^2 pixels x 8 samples, over the slab of depth 0.5 about
  z = 0.25 (it holds the whole lattice, whose particles are wrapped into the
  box) and over a thinner slab that cuts it: FP64 pixels and sample counts
  against the oracle on sampled 4-row bands, the mixed map against the FP64
  map everywhere;
* the same kernels (slab cut, 64-bit offsets, 4 lanes per pixel, both second
  pass forms) on a 131,072-particle lattice, forced by SDTFE_WIDE_RECORDS=1;
* 512 halo cutouts in one batched launch and a Poisson stream of 256 cutout
  requests through the server: every FP64 item against the oracle on one
  band of its map, mixed items against the FP64 ones.

Bars (the map's, src/dtfe.c:554-654): FP64 within 1e-10 relative per pixel,
sample counts and tets crossed exact; mixed within 1e-5 of FP64.
"""
import os
import sys
import time

import numpy as np
import pytest

from sdtfe_amd import capi, serve, workloads
from sdtfe_amd.tessellate import to_particle_data, triangulate

pytestmark = pytest.mark.gpu

RTOL_FP64 = 1e-10
RTOL_MIXED = 1e-5
THREADS = min(16, os.cpu_count() or 1)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    if a.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def _oracle_field(oracle, pd, tetra):
    """Oracle density (into pd[:, 3]), gradients and forward-hull rows."""
    oracle.density(pd, tetra, 1.0)
    return oracle.gradients(tetra, pd), oracle.triangles(tetra, pd)


# ------------------------------------------------ slab cut + 64-bit offsets

@pytest.mark.parametrize("defer_form", ["wave", "sample"])
def test_wide_slab_kernels_on_lattice(oracle, defer_form):
    """map_mixed_kernel with a cut slab and 64-bit face-record offsets (the
    large map's first pass) and the 64-bit second pass, on a 64 x 64 x 32
    lattice: equal to the 32-bit kernels bit for bit, FP64 equal to the
    oracle, mixed within 1e-5."""
    _say(f"[gpu] wide records, cut slab, lattice 131072, second pass {defer_form}")
    pos = workloads.jittered_lattice(64, 64, 32, seed=6)
    tetra = triangulate(pos)
    G, S, seed, lz, cz = 256, 8, 17, 0.3, 0.25
    assert pos[:, 2].min() < cz - lz / 2 or pos[:, 2].max() > cz + lz / 2, "slab must cut"
    kw = dict(seed=seed, box_len_z=lz, center=(0.5, 0.5, cz))
    env = {"SDTFE_DEFER_WAVE_ENTRIES": "0" if defer_form == "sample" else "1000000000"}
    res = {}
    os.environ.update(env)
    try:
        with capi.Tessellation(to_particle_data(pos), tetra) as t:
            for wide in ("", "1"):
                os.environ["SDTFE_WIDE_RECORDS"] = wide
                res[wide] = t.map(G, S, want_navg=True, precision=capi.PREC_MIXED, **kw)
            os.environ.pop("SDTFE_WIDE_RECORDS", None)
            r64, s64, n64 = t.map(G, S, want_navg=True, **kw)
    finally:
        for k in list(env) + ["SDTFE_WIDE_RECORDS"]:
            os.environ.pop(k, None)
    (m0, st0, n0), (m1, st1, n1) = res[""], res["1"]
    np.testing.assert_array_equal(m1, m0)
    np.testing.assert_array_equal(n1, n0)
    assert st1["tet_steps"] == st0["tet_steps"] == s64["tet_steps"]
    assert st1["fp64_redo"] > 0, "no ray took the second pass"
    np.testing.assert_array_equal(n1, n64)
    assert _rel(m1, r64) <= RTOL_MIXED
    pd = to_particle_data(pos)
    nabla, tri = _oracle_field(oracle, pd, tetra)
    ref = np.zeros(G * G)
    navg = np.zeros(G * G, dtype=np.int32)
    oracle.map_bands(pd, tetra, nabla, tri, G, 1.0, lz, (0.5, 0.5, cz), S, 0.0, seed,
                     band_begin=0, band_step=1, n_threads=THREADS, rho=ref, n_avg=navg)
    np.testing.assert_array_equal(n64, navg)
    assert _rel(r64, ref) <= RTOL_FP64
    _say(f"  fp64 redo {st1['fp64_redo']}, mixed vs fp64 {_rel(m1, r64):.2e}")


# ------------------------------------------------------- the large map

@pytest.mark.timeout(900)
@pytest.mark.parametrize("lz,cz", [(0.5, 0.25), (0.2, 0.3)], ids=["whole", "cut"])
def test_large_lattice_map(oracle, large_lattice, lz, cz):
    """8,388,608 particles, 4096^2 x 8, slab lz deep about z = cz."""
    pos, tetra = large_lattice
    _say(f"[gpu] large map: 8,388,608-particle lattice, 4096^2 x 8, slab {lz} about {cz}")
    G, S, seed = 4096, 8, 0
    full = pos[:, 2].min() >= cz - lz / 2 and pos[:, 2].max() <= cz + lz / 2
    assert full == (lz == 0.5)
    kw = dict(seed=seed, box_len_z=lz, center=(0.5, 0.5, cz))
    with capi.Tessellation(to_particle_data(pos), tetra) as t:
        r64, s64, n64 = t.map(G, S, want_navg=True, **kw)
        r32, s32, n32 = t.map(G, S, want_navg=True, precision=capi.PREC_MIXED, **kw)
    _say(f"  fp64 walk {s64['t_walk_ms']:.1f} ms, mixed {s32['t_walk_ms']:.1f} ms, "
         f"handed over {s32['fp64_redo']}")
    assert s64["samples"] == G * G * S and s64["step_cap"] == 0
    np.testing.assert_array_equal(n32, n64)
    assert s32["tet_steps"] == s64["tet_steps"]
    e = _rel(r32, r64)
    _say(f"  mixed vs fp64 max rel err {e:.2e}")
    assert e <= RTOL_MIXED
    del r32, n32
    pd = to_particle_data(pos)
    t0 = time.perf_counter()
    nabla, tri = _oracle_field(oracle, pd, tetra)
    bands = G // 4
    step = bands // 6  # six bands spread over the map
    ref = np.zeros(G * G)
    navg = np.zeros(G * G, dtype=np.int32)
    _, rays, _ = oracle.map_bands(pd, tetra, nabla, tri, G, 1.0, lz, (0.5, 0.5, cz), S, 0.0,
                                  seed, band_begin=step // 2, band_step=step,
                                  n_threads=THREADS, rho=ref, n_avg=navg)
    sel = oracle.band_rows(G, step // 2, step)
    _say(f"  oracle: {rays} sample-rays on {sel.sum() // (4 * G)} bands in "
         f"{time.perf_counter() - t0:.1f} s")
    np.testing.assert_array_equal(n64[sel], navg[sel])
    e = _rel(r64[sel], ref[sel])
    _say(f"  fp64 vs oracle max rel err {e:.2e}, bit-equal {(r64[sel] == ref[sel]).mean():.4f}")
    assert e <= RTOL_FP64


@pytest.mark.timeout(900)
def test_clustered_1m_map(oracle):
    """1,048,576 clustered particles (30% uniform, 70% in 200 halos), 2048^2
    x 16 over the whole box: FP64 pixels and sample counts against the
    oracle on eight sampled 4-row bands, the mixed map against the FP64 map
    everywhere with the same tets crossed."""
    _say("[gpu] clustered 1,048,576 particles, 2048^2 x 16")
    t0 = time.perf_counter()
    pos = workloads.clustered(1 << 20, seed=0)
    tetra = triangulate(pos)
    _say(f"  {tetra.shape[0]} tets in {time.perf_counter() - t0:.1f} s")
    G, S, seed = 2048, 16, 0
    with capi.Tessellation(to_particle_data(pos), tetra) as t:
        r64, s64, n64 = t.map(G, S, want_navg=True, seed=seed)
        r32, s32, n32 = t.map(G, S, want_navg=True, seed=seed, precision=capi.PREC_MIXED)
    _say(f"  fp64 walk {s64['t_walk_ms']:.1f} ms, mixed {s32['t_walk_ms']:.1f} ms, "
         f"handed over {s32['fp64_redo']}")
    assert s64["samples"] == G * G * S and s64["step_cap"] == 0
    np.testing.assert_array_equal(n32, n64)
    assert s32["tet_steps"] == s64["tet_steps"]
    e = _rel(r32, r64)
    _say(f"  mixed vs fp64 max rel err {e:.2e}")
    assert e <= RTOL_MIXED
    pd = to_particle_data(pos)
    nabla, tri = _oracle_field(oracle, pd, tetra)
    step = (G // 4) // 8
    ref = np.zeros(G * G)
    navg = np.zeros(G * G, dtype=np.int32)
    _, rays, _ = oracle.map_bands(pd, tetra, nabla, tri, G, 1.0, 1.0, (0.5, 0.5, 0.5), S, 0.0,
                                  seed, band_begin=step // 2, band_step=step,
                                  n_threads=THREADS, rho=ref, n_avg=navg)
    sel = oracle.band_rows(G, step // 2, step)
    np.testing.assert_array_equal(n64[sel], navg[sel])
    e = _rel(r64[sel], ref[sel])
    _say(f"  oracle: {rays} sample-rays; fp64 vs oracle max rel err {e:.2e}, bit-equal "
         f"{(r64[sel] == ref[sel]).mean():.4f}")
    assert e <= RTOL_FP64


@pytest.fixture(scope="module")
def large_lattice():
    from sdtfe_amd.tessellate_chunked import triangulate_chunked
    t0 = time.perf_counter()
    pos = workloads.jittered_lattice(seed=0)
    tetra = triangulate_chunked(pos, slabs=16, workers=THREADS, gpu=True)
    _say(f"  large lattice: {tetra.shape[0]} tets in {time.perf_counter() - t0:.1f} s")
    assert tetra.shape[0] > (1 << 25), "the large map must take the 64-bit face-record offsets"
    return pos, tetra


# ------------------------------------------- 512 cutouts, 256 requests

@pytest.fixture(scope="module")
def cutouts512():
    t0 = time.perf_counter()
    reqs = serve.make_requests(512, seed=0, workers=THREADS)
    _say(f"  512 cutouts tessellated in {time.perf_counter() - t0:.1f} s: "
         f"{sum(r.pd.shape[0] for r in reqs)} particles, "
         f"{sum(r.tetra.shape[0] for r in reqs)} tets")
    return reqs


@pytest.fixture(scope="module")
def batch_maps(cutouts512):
    """Both forms of all 512 maps from one batched launch each; lanes per
    pixel cycle through 1, 2, 4, 8, 16 over the items."""
    reqs = cutouts512
    tess = [capi.Tessellation(r.pd, r.tetra) for r in reqs]
    lpp = [(1, 2, 4, 8, 16)[i % 5] for i in range(len(reqs))]
    out = {}
    try:
        for prec in (capi.PREC_FP64, capi.PREC_MIXED):
            bufs = [capi.DeviceBuffer(r.grid * r.grid * 8) for r in reqs]
            for b in bufs:
                b.zero()
            with capi.Batch() as b:
                items = [(t, t.params(r.grid, r.samples, seed=r.seed, precision=prec), buf)
                         for t, r, buf in zip(tess, reqs, bufs)]
                st, tot = b.run(items, lanes_per_pixel=lpp)
            out[prec] = ([buf.to_host() for buf in bufs], st, tot)
            for buf in bufs:
                buf.close()
            _say(f"  batch of 512, precision {prec}: {tot['t_walk_ms']:.1f} ms")
    finally:
        for t in tess:
            t.close()
    return out


@pytest.fixture(scope="module")
def oracle_bands(cutouts512, oracle):
    """The oracle's pixels on one 4-row band of every cutout's map (band i mod
    bands), with their sample counts."""
    res = []
    t0 = time.perf_counter()
    for i, r in enumerate(cutouts512):
        pd = r.pd.copy()
        nabla, tri = _oracle_field(oracle, pd, r.tetra)
        G = r.grid
        band = i % (G // 4)
        ref = np.zeros(G * G)
        navg = np.zeros(G * G, dtype=np.int32)
        oracle.map_bands(pd, r.tetra, nabla, tri, G, 1.0, 1.0, (0.5, 0.5, 0.5), r.samples, 0.0,
                         r.seed, band_begin=band, band_step=G // 4, n_threads=THREADS, rho=ref,
                         n_avg=navg)
        res.append((oracle.band_rows(G, band, G // 4), ref, navg))
    _say(f"  oracle bands of 512 cutouts in {time.perf_counter() - t0:.1f} s")
    return res


@pytest.mark.timeout(900)
def test_batch_512_cutouts_match_oracle(cutouts512, batch_maps, oracle_bands):
    _say("[gpu] 512 cutouts in one batched launch against the oracle")
    maps64, st64, _ = batch_maps[capi.PREC_FP64]
    maps32, st32, _ = batch_maps[capi.PREC_MIXED]
    worst64 = worst32 = 0.0
    for r, m64, m32, a, b, (sel, ref, navg) in zip(cutouts512, maps64, maps32, st64, st32,
                                                    oracle_bands):
        assert a["samples"] == r.rays and b["samples"] == r.rays
        assert a["tet_steps"] == b["tet_steps"]
        worst64 = max(worst64, _rel(m64[sel], ref[sel]))
        worst32 = max(worst32, _rel(m32, m64))
    _say(f"  fp64 vs oracle {worst64:.2e}, mixed vs fp64 {worst32:.2e}")
    assert worst64 <= RTOL_FP64
    assert worst32 <= RTOL_MIXED


@pytest.mark.timeout(900)
@pytest.mark.parametrize("prec", [capi.PREC_FP64, capi.PREC_MIXED])
def test_serve_256_requests_match_oracle(cutouts512, batch_maps, oracle_bands, prec):
    """A Poisson stream of 256 cutout requests (150 per second) through the
    continuous-batching server: FP64 maps equal to the batched ones bit for
    bit and to the oracle on their band, mixed maps within 1e-5 of FP64."""
    _say(f"[gpu] server: 256 requests at 150/s, precision {prec}")
    reqs = cutouts512[:256]
    maps = []
    with capi.Server(max_active=64, tile_budget=4096, precision=prec) as sv:
        res = serve.run_stream(sv, reqs, rate_hz=150.0, seed=2, maps=maps)
    _say(f"  p50 {res['p50_ms']:.1f} ms, p99 {res['p99_ms']:.1f} ms, "
         f"{res['sample_rays_per_s']:.3e} sample-rays/s")
    maps64 = batch_maps[capi.PREC_FP64][0]
    worst = 0.0
    for m, m64, (sel, ref, _) in zip(maps, maps64, oracle_bands):
        if prec == capi.PREC_FP64:
            np.testing.assert_array_equal(m, m64)
            worst = max(worst, _rel(m[sel], ref[sel]))
        else:
            worst = max(worst, _rel(m, m64))
    _say(f"  worst rel err {worst:.2e}")
    assert worst <= (RTOL_FP64 if prec == capi.PREC_FP64 else RTOL_MIXED)
    assert res["p99_ms"] >= res["p50_ms"] > 0
