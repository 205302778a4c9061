"""The mixed-precision walk against the FP64 walk on the GPU.

Same topology by construction (every side sign equals the FP64 walk's), so
tets crossed, perturbations and n_avg must be identical; values carry only
the FP32 integrand's rounding.  Tolerance: 1e-5 relative per pixel (and per
ray), the bound the mixed form is held to.
"""
import os
import sys

import numpy as np
import pytest

from sdtfe_amd import capi, workloads
from sdtfe_amd.tessellate import to_particle_data, triangulate

pytestmark = pytest.mark.gpu

RTOL_MIXED = 1e-5


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _relerr(a, b):
    m = np.abs(b) > 0
    return np.abs(a - b)[m] / np.abs(b)[m]


def test_mixed_rays_match_fp64(golden):
    _say("[gpu] mixed rays g2")
    g = golden("g2_rays.npz")
    with capi.Tessellation(to_particle_data(g["pos"]), g["tetra"]) as t:
        args = (g["x"], g["y"], g["pixel"], g["sample"], float(g["off_z"]), float(g["lz"]),
                float(g["dxy"]), int(g["seed"]))
        v64, c64, s64 = t.trace(*args, entry=g["entry"])
        v32, c32, s32 = t.trace(*args, entry=g["entry"], precision=capi.PREC_MIXED)
    np.testing.assert_array_equal(c32, c64)
    assert s32["perturb_steps"] == s64["perturb_steps"]
    e = _relerr(v32, v64)
    _say(f"  max rel err {e.max():.2e}, fp64 redo {s32['fp64_redo']}")
    assert e.max() <= RTOL_MIXED


@pytest.mark.parametrize("kind,n,G,S,box_z,cz", [
    ("uniform", 40000, 256, 4, 1.0, 0.5),
    ("clustered", 131072, 256, 8, 1.0, 0.5),
    ("clustered-slab", 131072, 192, 4, 0.3, 0.45),
    # 4.2M sample-rays: rare near-edge cases (sides the FP32 filter leaves
    # open, on the first tet's faces too) occur many times
    ("clustered-large", 262144, 512, 16, 1.0, 0.5),
    # the large-map layout at 1/512 of its size: a 32x32x16 lattice with
    # Gaussian displacements (0.3 spacing) and 5% of the particles in halos
    ("lattice", 32 * 32 * 16, 128, 8, 0.5, 0.25),
])
def test_mixed_map_matches_fp64(kind, n, G, S, box_z, cz):
    _say(f"[gpu] mixed map {kind}")
    if kind == "uniform":
        pos = workloads.uniform(n, 7)
    elif kind == "lattice":
        pos = workloads.jittered_lattice(32, 32, 16, seed=5)
    else:
        pos = workloads.clustered(n, seed=5)
    tetra = triangulate(pos)
    kw = dict(seed=13, box_len_z=box_z, center=(0.5, 0.5, cz))
    with capi.Tessellation(to_particle_data(pos), tetra) as t:
        r64, s64, n64 = t.map(G, S, want_navg=True, **kw)
        r32, s32, n32 = t.map(G, S, want_navg=True, precision=capi.PREC_MIXED, **kw)
    np.testing.assert_array_equal(n32, n64)
    assert s32["tet_steps"] == s64["tet_steps"]
    assert s32["perturb_steps"] == s64["perturb_steps"]
    e = _relerr(r32, r64)
    _say(f"  max rel err {e.max():.2e}, median {np.median(e):.2e}, "
         f"fp64 redo / step {s32['fp64_redo'] / max(1, s32['tet_steps']):.2e}")
    assert e.max() <= RTOL_MIXED


@pytest.mark.parametrize("S,box_z,cz", [(8, 1.0, 0.5), (16, 1.0, 0.5), (4, 0.3, 0.45)])
def test_zero_tolerance_gives_the_fp64_map(S, box_z, cz):
    """With the per-ray tolerance at 0 no FP32 value can stand: a ray handed
    over at its end (its bound) and a ray handed over in mid-ray (a side the
    filter leaves open, whose mixed prefix must then meet the tolerance too)
    both go to the FP64 walk from their entry, so the mixed map equals the
    FP64 map bit for bit, sample counts and tets crossed included."""
    _say(f"[gpu] mixed map at tolerance 0, {S} samples, slab {box_z}")
    pos = workloads.clustered(65536, seed=11)
    tetra = triangulate(pos)
    kw = dict(seed=4, box_len_z=box_z, center=(0.5, 0.5, cz))
    with capi.Tessellation(to_particle_data(pos), tetra) as t:
        r64, s64, n64 = t.map(256, S, want_navg=True, **kw)
        os.environ["SDTFE_MIXED_TOL"] = "0"
        try:
            r32, s32, n32 = t.map(256, S, want_navg=True, precision=capi.PREC_MIXED, **kw)
        finally:
            os.environ.pop("SDTFE_MIXED_TOL", None)
    _say(f"  handed over {s32['fp64_redo']} of {s32['samples']}")
    assert s32["fp64_redo"] > 0.5 * s32["samples"]
    np.testing.assert_array_equal(n32, n64)
    assert s32["tet_steps"] == s64["tet_steps"]
    np.testing.assert_array_equal(r32, r64)


def test_mixed_map_is_deterministic():
    """Two runs of a mixed map agree bit for bit: each pixel has one writer
    and a fixed summation order, also where samples go to the FP64 pass."""
    _say("[gpu] mixed map twice")
    pos = workloads.clustered(131072, seed=9)
    tetra = triangulate(pos)
    with capi.Tessellation(to_particle_data(pos), tetra) as t:
        a, sa, na = t.map(256, 8, want_navg=True, seed=3, precision=capi.PREC_MIXED)
        b, sb, nb = t.map(256, 8, want_navg=True, seed=3, precision=capi.PREC_MIXED)
    _say(f"  handed to fp64: {sa['fp64_redo']} of {sa['samples']} samples")
    assert sa["fp64_redo"] > 0  # the second pass ran
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(na, nb)
    counts = lambda st: {k: v for k, v in st.items() if not k.startswith("t_")}  # noqa: E731
    assert counts(sa) == counts(sb)


def test_second_pass_forms_agree():
    """The second pass as a wave per queued pixel and as a lane per queued
    sample (the form for long queues, forced here by a limit of 0 queued
    pixels) gives the same map, n_avg and counters bit for bit."""
    _say("[gpu] mixed map, second pass per pixel vs per sample")
    pos = workloads.clustered(262144, seed=5)
    tetra = triangulate(pos)
    kw = dict(seed=13, want_navg=True, precision=capi.PREC_MIXED)
    with capi.Tessellation(to_particle_data(pos), tetra) as t:
        res = {}
        for lim in ("", "0"):
            os.environ["SDTFE_DEFER_WAVE_ENTRIES"] = lim
            try:
                res[lim] = t.map(512, 16, **kw)
            finally:
                os.environ.pop("SDTFE_DEFER_WAVE_ENTRIES", None)
        r64, _ = t.map(512, 16, seed=13)
    (a, sa, na), (b, sb, nb) = res[""], res["0"]
    _say(f"  handed to fp64: {sa['fp64_redo']} of {sa['samples']} samples")
    assert sa["fp64_redo"] > 0
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(na, nb)
    counts = lambda st: {k: v for k, v in st.items() if not k.startswith("t_")}  # noqa: E731
    assert counts(sa) == counts(sb)
    assert _relerr(b, r64).max() <= RTOL_MIXED
