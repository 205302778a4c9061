"""Serving workload: a Poisson stream of cutout-map requests against the
continuous-batching server (sdtfe_server), reporting p50/p99 map latency and
sample-rays/s.

Requests are drawn like the batched-cutout workload (particle counts
log-uniform, grids log-uniform and rounded to a multiple of 4, 4-16 samples
per pixel); each carries its particles and host (Qhull) tessellation, which
are made before the stream starts.
"""
from __future__ import annotations

import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass

import numpy as np

from . import capi, workloads
from .tessellate import to_particle_data, triangulate


@dataclass
class Request:
    pd: np.ndarray
    tetra: np.ndarray
    grid: int
    samples: int
    seed: int

    @property
    def rays(self) -> int:
        return self.grid * self.grid * self.samples


def _one(spec: workloads.CutoutSpec) -> Request:
    pos = workloads.cutout(spec.n_particles, spec.seed)
    return Request(to_particle_data(pos), triangulate(pos), spec.map.grid_dim,
                   spec.map.n_samples, spec.seed)


def make_requests(n: int, seed: int = 0, n_hi: float = 2e5, g_hi: float = 1024,
                  workers: int = 8) -> list[Request]:
    specs = workloads.batch_specs(n, seed=seed, n_hi=n_hi, g_hi=g_hi)
    with ThreadPoolExecutor(max_workers=workers) as ex:  # scipy's Qhull releases the GIL
        return list(ex.map(_one, specs))


def params_for(server: capi.Server, req: Request) -> capi.MapParams:
    p = capi.MapParams()
    p.grid_dim, p.n_mc_samp = req.grid, req.samples
    p.box_len, p.box_len_z = 1.0, 1.0
    p.center[:] = [0.5, 0.5, 0.5]
    p.delta_sample, p.seed = 0.0, req.seed
    p.precision, p.max_steps = server.precision, 0
    return p


def run_stream(server: capi.Server, reqs: list[Request], rate_hz: float | None,
               seed: int = 0, maps: list | None = None) -> dict:
    """Submit reqs at Poisson arrival times (rate_hz; None = all at once),
    wait for all, and summarise latencies and throughput.  maps (a list,
    optional) receives each request's map in request order."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_hz, len(reqs)) if rate_hz else np.zeros(len(reqs))
    ids: list[int] = []
    results: dict[int, dict] = {}
    errors: list[str] = []
    got: dict[int, np.ndarray] = {}
    done = threading.Event()

    def waiter():
        # waits for every request, also past a failed one, so the server drains
        try:
            k = 0
            while k < len(reqs):
                while k >= len(ids) and not errors:
                    time.sleep(0.0005)
                if k >= len(ids):
                    break  # submission stopped on an error
                try:
                    rho, info = server.wait(ids[k])
                    results[k] = info
                    if maps is not None:
                        got[k] = rho
                except capi.SdtfeError as e:
                    errors.append(str(e))
                k += 1
        finally:
            done.set()

    th = threading.Thread(target=waiter, daemon=True)
    t0 = time.perf_counter()
    th.start()
    t_next = t0
    try:
        for r, gap in zip(reqs, gaps):
            t_next += gap
            dt = t_next - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            ids.append(server.submit(r.pd, r.tetra, params_for(server, r)))
    except capi.SdtfeError as e:
        errors.append(str(e))
    done.wait()
    if errors:
        raise RuntimeError(f"{len(errors)} request(s) failed; first: {errors[0]}")
    wall = time.perf_counter() - t0
    if maps is not None:
        maps.extend(got[k] for k in range(len(reqs)))
    lat = np.array([results[k]["latency_ms"] for k in range(len(reqs))])
    rays = sum(r.rays for r in reqs)
    return dict(requests=len(reqs), wall_s=wall, rate_hz=rate_hz,
                p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                mean_ms=float(lat.mean()), sample_rays_per_s=rays / wall,
                queue_p99_ms=float(np.percentile([results[k]["queue_ms"] for k in results], 99)),
                server=server.info())
