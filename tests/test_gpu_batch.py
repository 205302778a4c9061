"""Batched persistent launch: every item's map equals its standalone map
(FP64: bit for bit, whatever the sample tiling), across uneven cutouts and
repeated graph replays."""
import os
import sys

import numpy as np
import pytest

from sdtfe_amd import autotune, capi, workloads
from sdtfe_amd.tessellate import to_particle_data, triangulate

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def cutouts():
    specs = [(1500, 3, 64, 4), (6000, 4, 100, 7), (20000, 5, 48, 16), (3000, 6, 128, 5)]
    out = []
    for n, seed, G, S in specs:
        pos = workloads.cutout(n, seed)
        t = capi.Tessellation(to_particle_data(pos), triangulate(pos))
        out.append((t, G, S, seed))
    yield out
    for t, *_ in out:
        t.close()


@pytest.mark.parametrize("prec", [capi.PREC_FP64, capi.PREC_MIXED])
@pytest.mark.parametrize("lpp", [1, 4])
def test_batch_equals_standalone(cutouts, prec, lpp):
    _say(f"[gpu] batch prec={prec} lanes/pixel={lpp}")
    alone = [t.map(G, S, seed=seed, precision=prec)[0] for t, G, S, seed in cutouts]
    bufs = [capi.DeviceBuffer(G * G * 8) for _, G, _, _ in cutouts]
    with capi.Batch() as b:
        for rep in range(2):  # second run replays the captured graph
            for buf in bufs:
                buf.zero()
            items = [(t, t.params(G, S, seed=seed, precision=prec), buf)
                     for (t, G, S, seed), buf in zip(cutouts, bufs)]
            st, tot = b.run(items, lanes_per_pixel=[lpp] * len(items))
            for (t, G, S, seed), buf, ref, s in zip(cutouts, bufs, alone, st):
                got = buf.to_host()
                # same walk, same summation order (the mixed form's handed-over
                # samples are added in sample order by its second pass)
                np.testing.assert_array_equal(got, ref)
                assert s["samples"] == G * G * S
    for buf in bufs:
        buf.close()
    assert tot["t_walk_ms"] > 0


def test_autotuner_picks_a_candidate(cutouts, tmp_path):
    _say("[gpu] autotuner")
    t, G, S, seed = cutouts[2]
    tuner = autotune.Autotuner(str(tmp_path / "tune.json"), candidates=(1, 4), repeats=1)
    with capi.Batch() as b:
        lpp = tuner.lanes_for(b, t, t.params(G, S, seed=seed))
    assert lpp in (1, 4)
    tuner.save()
    again = autotune.Autotuner(str(tmp_path / "tune.json"))
    assert again.table == tuner.table


def test_wide_record_kernels_agree(cutouts):
    """The kernels with 64-bit face-record offsets (taken for tessellations
    beyond 2^25 tets; SDTFE_WIDE_RECORDS=1 takes them for any size) give the
    same mixed maps as the 32-bit ones, single maps and batches alike."""
    _say("[gpu] wide-record kernels vs narrow, single maps and a batch")
    res = {}
    for wide in ("", "1"):
        os.environ["SDTFE_WIDE_RECORDS"] = wide
        try:
            alone = [t.map(G, S, seed=seed, precision=capi.PREC_MIXED)
                     for t, G, S, seed in cutouts]
            bufs = [capi.DeviceBuffer(G * G * 8) for _, G, _, _ in cutouts]
            for buf in bufs:
                buf.zero()
            with capi.Batch() as b:
                items = [(t, t.params(G, S, seed=seed, precision=capi.PREC_MIXED), buf)
                         for (t, G, S, seed), buf in zip(cutouts, bufs)]
                st, _ = b.run(items, lanes_per_pixel=[4] * len(items))
            res[wide] = (alone, [buf.to_host() for buf in bufs], st)
            for buf in bufs:
                buf.close()
        finally:
            os.environ.pop("SDTFE_WIDE_RECORDS", None)
    (a0, b0, s0), (a1, b1, s1) = res[""], res["1"]
    for (m0, st0), (m1, st1), x0, x1, bs0, bs1 in zip(a0, a1, b0, b1, s0, s1):
        np.testing.assert_array_equal(m1, m0)
        np.testing.assert_array_equal(x1, x0)
        np.testing.assert_array_equal(x1, m0)
        assert st1["tet_steps"] == st0["tet_steps"] and bs1["tet_steps"] == bs0["tet_steps"]
