"""Continuous-batching server: every request's map equals the same map
computed alone (FP64: bit for bit), with requests arriving while others run,
under a tight HBM budget that forces admission to wait."""
import sys
import threading
import time

import numpy as np
import pytest

from sdtfe_amd import capi, serve, workloads

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def requests():
    return serve.make_requests(10, seed=3, n_hi=3e4, g_hi=192)


@pytest.mark.parametrize("prec", [capi.PREC_FP64, capi.PREC_MIXED])
def test_mix_equals_alone(requests, prec):
    _say(f"[gpu] server mix vs alone, precision {prec}")
    alone = []
    for r in requests:
        with capi.Tessellation(r.pd, r.tetra) as t:
            alone.append(t.map(r.grid, r.samples, seed=r.seed, precision=prec)[0])
    with capi.Server(max_active=4, tile_budget=64, precision=prec) as sv:
        ids = []
        for k, r in enumerate(requests):
            ids.append(sv.submit(r.pd, r.tetra, serve.params_for(sv, r)))
            if k % 3 == 2:
                time.sleep(0.01)  # later requests arrive while earlier ones run
        outs = [sv.wait(i) for i in ids]
        info = sv.info()
    for (got, inf), ref, r in zip(outs, alone, requests):
        assert inf["status"] == 0
        assert inf["stats"]["samples"] == r.rays
        if prec == capi.PREC_FP64:
            np.testing.assert_array_equal(got, ref)
        else:
            np.testing.assert_allclose(got, ref, rtol=1e-5, atol=0)
    assert info["launches"] >= 2


def test_admission_respects_budget(requests):
    _say("[gpu] server admission under a tight budget")
    biggest = max(r.tetra.shape[0] for r in requests)
    budget = 304 * biggest + (1 << 22)  # roughly one request at a time
    peak = [0]
    with capi.Server(hbm_budget_bytes=budget, max_active=8, tile_budget=32) as sv:
        stop = threading.Event()

        def watch():
            while not stop.is_set():
                peak[0] = max(peak[0], sv.info()["used_bytes"])
                time.sleep(0.001)

        th = threading.Thread(target=watch)
        th.start()
        ids = [sv.submit(r.pd, r.tetra, serve.params_for(sv, r)) for r in requests[:5]]
        infos = [sv.wait(i)[1] for i in ids]
        stop.set()
        th.join()
    assert all(i["status"] == 0 for i in infos)
    one = max(i["footprint_bytes"] for i in infos)
    assert peak[0] <= max(budget, one)


def test_poisson_stream_reports_latency(requests):
    _say("[gpu] poisson stream")
    with capi.Server(max_active=8, tile_budget=256) as sv:
        res = serve.run_stream(sv, requests, rate_hz=50.0, seed=1)
    assert res["p50_ms"] > 0 and res["p99_ms"] >= res["p50_ms"]
    assert res["sample_rays_per_s"] > 0


def test_failed_request_reports_its_error(requests):
    """A request whose tetra_data names a vertex past the particles fails at
    set-up (the index check, before any walk); its waiter gets the message,
    and the requests around it still complete."""
    _say("[gpu] server: a failing request among good ones")
    good = requests[:3]
    bad_tets = good[1].tetra.copy()
    bad_tets[5, 2] = good[1].pd.shape[0] + 3
    with capi.Server(max_active=4, tile_budget=64) as sv:
        ids = [sv.submit(good[0].pd, good[0].tetra, serve.params_for(sv, good[0])),
               sv.submit(good[1].pd, bad_tets, serve.params_for(sv, good[1])),
               sv.submit(good[2].pd, good[2].tetra, serve.params_for(sv, good[2]))]
        _, i0 = sv.wait(ids[0])
        with pytest.raises(capi.SdtfeError, match="out of range"):
            sv.wait(ids[1])
        _, i2 = sv.wait(ids[2])
    assert i0["status"] == 0 and i2["status"] == 0


def test_close_drains_unwaited_requests(requests):
    """Closing a server with requests nobody waited for drains it: every map
    is walked, copied out into the caller's array and freed before close
    returns, and equals the same map computed alone (FP64: bit for bit)."""
    _say("[gpu] server: close with requests never waited for")
    reqs = requests[:6]
    alone = []
    for r in reqs:
        with capi.Tessellation(r.pd, r.tetra) as t:
            alone.append(t.map(r.grid, r.samples, seed=r.seed)[0])
    outs = [np.zeros(r.grid * r.grid) for r in reqs]
    sv = capi.Server(max_active=4, tile_budget=64)
    for r, o in zip(reqs, outs):
        sv.submit(r.pd, r.tetra, serve.params_for(sv, r), rho=o)
    sv.close()
    for got, ref in zip(outs, alone):
        np.testing.assert_array_equal(got, ref)
