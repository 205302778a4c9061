"""sdtfe_map_device_noalloc (the map entry point that allocates nothing)
through the C ABI against sdtfe_map_device on the same tessellation and
parameters: the same map bit for bit (both forms), two calls in a row with the
second smaller, both second-pass forms of the mixed map and the
wave-per-pixel finish of a long queue, and the size limit."""
import os
import sys

import numpy as np
import pytest

from sdtfe_amd import capi, workloads
from sdtfe_amd.tessellate import to_particle_data, triangulate

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def tess():
    pos = workloads.clustered(65536, seed=3)
    t = capi.Tessellation(to_particle_data(pos), triangulate(pos))
    yield t
    t.close()


def _map(t, fn, G, S, prec, seed, start):
    buf = capi.DeviceBuffer(G * G * 8)
    capi.check(capi.lib().sdtfe_memcpy_h2d(buf.ptr, start.ctypes.data, start.nbytes))
    st = fn(t.params(G, S, seed=seed, precision=prec), buf)
    out = buf.to_host()
    buf.close()
    return out, st


# the FP64 map has no second pass: one form; the mixed map: all three
@pytest.mark.parametrize("prec,form", [(capi.PREC_FP64, "wave"), (capi.PREC_MIXED, "wave"),
                                       (capi.PREC_MIXED, "sample"),
                                       (capi.PREC_MIXED, "sample-capped")])
def test_noalloc_equals_map_device(tess, prec, form):
    env = {"wave": {"SDTFE_DEFER_WAVE_ENTRIES": "1000000000"},
           "sample": {"SDTFE_DEFER_WAVE_ENTRIES": "0"},
           "sample-capped": {"SDTFE_DEFER_WAVE_ENTRIES": "0", "SDTFE_NOALLOC_SAMPLES": "4"}}[form]
    os.environ.update(env)
    try:
        # two calls in a row, the second with a smaller map; both accumulate
        # into the same non-zero start
        for k, (G, S, seed) in enumerate([(512, 16, 5), (132, 6, 9)]):
            _say(f"[gpu] no-alloc map {G}^2 x {S}, precision {prec}, second pass {form}")
            start = np.random.default_rng(k).random(G * G)
            want, st_w = _map(tess, tess.map_device, G, S, prec, seed, start)
            got, st_g = _map(tess, tess.map_device_noalloc, G, S, prec, seed, start)
            np.testing.assert_array_equal(got, want)
            for key in ("samples", "missed", "failed", "tet_steps", "perturb_steps",
                        "fp64_redo"):
                assert st_g[key] == st_w[key], key
            assert st_g["samples"] == G * G * S
            if prec == capi.PREC_MIXED and k == 0:
                assert st_g["fp64_redo"] > 4, "the second pass must have work"
    finally:
        for key in env:
            os.environ.pop(key, None)


def test_noalloc_rejects_oversized_mixed_map(tess):
    _say("[gpu] no-alloc size limit")
    buf = capi.DeviceBuffer(8)
    with pytest.raises(capi.SdtfeError, match="at most 1048576 pixels"):
        tess.map_device_noalloc(tess.params(1028, 4, precision=capi.PREC_MIXED), buf)
    buf.close()
