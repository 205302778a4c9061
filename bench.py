"""Surface-density map throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--config clustered|uniform|lattice|cutouts|serve|volume]

A step is one map of the workload with the tessellation resident in HBM:
Monte-Carlo sample-rays are placed, located on the forward hull and walked
through the tets, and pixel values are accumulated in HBM.  Default workload:
1,048,576 clustered particles (30% uniform + 200 NFW-like halos, seed 0),
2048x2048 pixels, 16 samples per pixel.  Particles are synthetic; the
tessellation (host Qhull) and the upload are outside the timed region.

Prints one JSON line: sample-rays/s of the whole job (value); the walk
kernel's roofline (physical HBM bytes per launch from the committed rocprofv3
FETCH_SIZE / WRITE_SIZE profile, stamped with the kernel's sources, over the
kernel time from HIP events on the library's stream in this run; the
algorithmic rate at 88 B per tet crossed beside it); a serving leg (a
Poisson stream of 256 cutout requests at 150 per second through the
continuous-batching server, both precisions: p50/p99 latency, capacity and
the mixed maps' error against the FP64 maps); and the CPU oracle timed on the
host cores over a bounded subset of the same map (every k-th 4-row band),
with the CPU model.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic bytes per tet crossed: the tet row (8 x int32 = 32 B), one new
# vertex position (3 x FP64 = 24 B; the other three are shared with the
# previous tet) and the field (rho_A + gradient, 4 x FP64 = 32 B).
BYTES_PER_STEP = 88
HBM_PEAK_GBS = 8000.0

CONFIGS = {
    "clustered": dict(n=1_048_576, grid=2048, samples=16, box_len=1.0, box_len_z=1.0,
                      center=(0.5, 0.5, 0.5)),
    "uniform": dict(n=32768, grid=256, samples=4, box_len=1.0, box_len_z=1.0,
                    center=(0.5, 0.5, 0.5)),
    "lattice": dict(n=8_388_608, grid=4096, samples=8, box_len=1.0, box_len_z=0.5,
                    center=(0.5, 0.5, 0.25)),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def particles(name):
    from sdtfe_amd import workloads
    if name == "clustered":
        return workloads.clustered(CONFIGS[name]["n"], seed=0)
    if name == "uniform":
        return workloads.uniform(CONFIGS[name]["n"], seed=0)
    return workloads.jittered_lattice(seed=0)


def rel_err(a, b):
    m = np.abs(b) > 0
    d = np.abs(a - b)[m] / np.abs(b)[m]
    return d


def time_form(tess, cfg, precision, steps, warmup, seed):
    from sdtfe_amd import capi
    G, S = cfg["grid"], cfg["samples"]
    buf = capi.DeviceBuffer(G * G * 8)
    p = tess.params(G, S, box_len=cfg["box_len"], box_len_z=cfg["box_len_z"],
                    center=cfg["center"], seed=seed, precision=precision)
    buf.zero()
    for _ in range(warmup):
        tess.map_device(p, buf, stats=True)
    buf.zero()
    tess.sync()
    walk_ms, kernel_ms, steps_per_map, st_last = [], [], [], None
    t0 = time.perf_counter()
    for _ in range(steps):
        st = tess.map_device(p, buf, stats=True)
        walk_ms.append(st["t_walk_ms"])
        kernel_ms.append(st["t_kernel_ms"])
        steps_per_map.append(st["tet_steps"])
        st_last = st
    tess.sync()
    wall = time.perf_counter() - t0
    # one more map, untimed, into a clean buffer: the map compared across forms
    buf.zero()
    tess.map_device(p, buf, stats=False)
    tess.sync()
    acc = buf.to_host()
    buf.close()
    return dict(wall_s=wall, ms_per_step=1e3 * wall / steps, walk_ms=float(np.mean(walk_ms)),
                kernel_ms=float(np.mean(kernel_ms)), tet_steps=int(np.mean(steps_per_map)),
                stats=st_last, map=acc)


# Sources that decide a walk kernel's memory traffic: a profile taken from
# other sources no longer describes the kernel that runs.
KERNEL_SOURCES = {
    "map_mixed_kernel": ["walk_mixed.hip", "walk_mixed_core.h", "walk_defer.h", "walk_common.h",
                         "walk_fp64_core.h", "sdtfe_internal.h", "field.hip"],
    "map_fp64_kernel": ["walk_fp64.hip", "walk_fp64_face.h", "walk_fp64_core.h", "walk_common.h",
                        "sdtfe_internal.h", "field.hip"],
    "volume_rays_kernel+volume_fold_kernel": ["volume.hip", "walk_fp64_core.h", "walk_common.h",
                                              "sdtfe_internal.h", "field.hip"],
}


def source_stamp(kernel):
    """SHA-256 (first 16 hex digits) over the kernel's source files."""
    import hashlib
    h = hashlib.sha256()
    for name in KERNEL_SOURCES.get(kernel, []):
        with open(os.path.join(ROOT, "sdtfe_amd", "csrc", name), "rb") as f:
            h.update(name.encode() + b"\0" + f.read())
    return h.hexdigest()[:16]


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (profiles/traffic.json), each
    scaled by the same counter's reading on a copy of known size, with the
    source stamp the profile was taken at; None when that workload was not
    profiled.  'stale' is set when the kernel's sources have changed since."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        tr = json.load(f).get(kernel, {}).get(workload)
    if tr is not None:
        tr = dict(tr)
        tr["stale"] = tr.get("source_stamp") != source_stamp(kernel)
    return tr


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or None


def _ordinal(k: int) -> str:
    if k == 1:
        return "single"  # every band
    suffix = "th" if 10 <= k % 100 <= 20 else {1: "st", 2: "nd", 3: "rd"}.get(k % 10, "th")
    return f"{k}{suffix}"


def cpu_baseline(pos, tetra, cfg, seed, target_s, gpu_map):
    """Oracle (C, OpenMP) on every k-th 4-row band of the same map."""
    from oracle import oracle as O
    from sdtfe_amd.tessellate import to_particle_data
    if not os.path.exists(O.LIB_PATH):
        O.build()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    pd = to_particle_data(pos)
    O.density(pd, tetra, 1.0)
    nabla = O.gradients(tetra, pd)
    tri = O.triangles(tetra, pd)
    G, S = cfg["grid"], cfg["samples"]
    bands = G // 4
    args = (pd, tetra, nabla, tri, G, cfg["box_len"], cfg["box_len_z"], cfg["center"], S, 0.0,
            seed)
    # calibrate on one band in the middle, then take every k-th band
    t0 = time.perf_counter()
    O.map_bands(*args, band_begin=bands // 2, band_step=bands, n_threads=threads)
    t_band = max(time.perf_counter() - t0, 1e-6)
    k = max(1, int(np.ceil(bands * t_band / target_s)))
    rho = np.zeros(G * G)
    t0 = time.perf_counter()
    _, rays, st = O.map_bands(*args, band_begin=k // 2, band_step=k, n_threads=threads, rho=rho)
    t = time.perf_counter() - t0
    rows = np.zeros(G, dtype=bool)
    for b in range(k // 2, bands, k):
        rows[4 * b:4 * b + 4] = True
    sel = np.zeros((G, G), dtype=bool)
    sel[rows, :] = True
    sel = sel.ravel()
    err = rel_err(gpu_map[sel], rho[sel]) if gpu_map is not None else np.zeros(1)
    ref = reference_baseline(pos, tetra, cfg, seed, target_s / 2, bands, args, threads, gpu_map)
    return dict(value=rays / t, unit="sample-rays/s", cores=threads, kind="port", reference=ref,
                sample=(f"every {_ordinal(k)} 4-row band of the {G}x{G} map ({rays} sample-rays, "
                        f"{100.0 * rays / (G * G * S):.2f}% of the map), same tessellation "
                        f"and sample stream, {t:.1f} s"),
                seconds=t, steps_per_ray=st["tet_steps"] / max(1, st["samples"] - st["missed"]),
                gpu_vs_cpu_max_rel_err=float(err.max()) if err.size else 0.0)


def reference_baseline(pos, tetra, cfg, seed, target_s, bands, port_args, threads, gpu_map):
    """The reference's own compiled pt_loc_2d and shoot_3d_ray (oracle/_ref,
    built from the reference's sources; single-threaded, as its OpenMP loop
    serialises on rand()) over every k-th 4-row band of the same map, and
    the port over the same bands, for bit-for-bit comparisons with the port
    and with the GPU's FP64 map.  None when the reference build is absent."""
    from oracle import oracle as O
    from sdtfe_amd.tessellate import to_particle_data
    if not O.ref_available():
        return None
    G, S = cfg["grid"], cfg["samples"]
    rs = O.RefSetup(to_particle_data(pos), tetra)
    try:
        geo = (G, cfg["box_len"], cfg["box_len_z"], cfg["center"], S, 0.0, seed)
        t0 = time.perf_counter()
        rs.map_bands(*geo, band_begin=bands // 2, band_step=bands)
        t_band = max(time.perf_counter() - t0, 1e-6)
        k = max(1, int(np.ceil(bands * t_band / target_s)))
        t0 = time.perf_counter()
        got, rays = rs.map_bands(*geo, band_begin=k // 2, band_step=k)
        t = time.perf_counter() - t0
    finally:
        rs.close()
    want, _, _ = O.map_bands(*port_args, band_begin=k // 2, band_step=k, n_threads=threads)
    rows = np.zeros(G, dtype=bool)
    for b in range(k // 2, bands, k):
        rows[4 * b:4 * b + 4] = True
    sel = np.repeat(rows, G)
    return dict(value=rays / t, unit="sample-rays/s", cores=1, kind="reference",
                sample=(f"every {_ordinal(k)} 4-row band of the {G}x{G} map ({rays} sample-rays), "
                        f"{t:.1f} s"),
                equal_to_port=bool(np.array_equal(got, want)),
                equal_to_gpu_fp64=(None if gpu_map is None
                                   else bool(np.array_equal(got[sel], gpu_map[sel]))))


def tessellation(name, pos, cache):
    """Host Qhull tessellation, optionally cached in an .npz (e.g. under /tmp
    between profiling passes of one session).  Above 2M particles the
    Delaunay tetrahedralisation is assembled from x-slabs in parallel (one
    Qhull run over 8M particles takes many minutes on one core), with the
    orientation and neighbour slots assembled on the GPU."""
    from sdtfe_amd.tessellate import triangulate
    from sdtfe_amd.tessellate_chunked import triangulate_chunked
    if cache and os.path.exists(cache):
        with np.load(cache) as z:
            if z["n"] == pos.shape[0] and str(z["name"]) == name:
                return z["tetra"]
    if pos.shape[0] > 2_000_000:
        workers = min(16, os.cpu_count() or 1)
        tetra = triangulate_chunked(pos, slabs=16, workers=workers, log=log, gpu=True)
    else:
        tetra = triangulate(pos)
    if cache:
        np.savez(cache, tetra=tetra, n=pos.shape[0], name=name)
    return tetra


def copy_calibration(gib: float = 1.0, reps: int = 3):
    """A plain device copy of known size (torch), for scaling rocprofv3's
    FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    n = int(gib * (1 << 30)) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "reps": reps}))


def _precision(a):
    from sdtfe_amd import capi
    return capi.PREC_FP64 if a.precision == "fp64" else capi.PREC_MIXED


def bench_cutouts(a):
    """Batched cutouts: many small maps, each its own resident tessellation,
    in one persistent launch (sdtfe_batch) against the same maps one launch
    at a time.  Cutout i: particles log-uniform in [1e3, 2e5], grid
    log-uniform in [64, 1024] (multiple of 4), 4-16 samples per pixel."""
    import collections
    from sdtfe_amd import autotune, capi, serve
    prec = _precision(a)
    t0 = time.perf_counter()
    reqs = serve.make_requests(a.cutouts, seed=a.seed, workers=16)
    log(f"[bench] {len(reqs)} cutouts tessellated in {time.perf_counter() - t0:.1f} s, "
        f"{sum(r.pd.shape[0] for r in reqs)} particles, {sum(r.tetra.shape[0] for r in reqs)} tets")
    tess = [capi.Tessellation(r.pd, r.tetra) for r in reqs]
    bufs = [capi.DeviceBuffer(r.grid * r.grid * 8) for r in reqs]
    params = [t.params(r.grid, r.samples, seed=r.seed, precision=prec) for t, r in zip(tess, reqs)]
    rays = sum(r.rays for r in reqs)
    for t in tess:
        t.sync()
    with capi.Batch() as b:
        # lanes per pixel (rays per wave from one pixel) per size class of
        # map, timed once per class before the timed region
        t0 = time.perf_counter()
        tuner = autotune.Autotuner()
        lpp = [tuner.lanes_for(b, t, p) for t, p in zip(tess, params)]
        tune_s = time.perf_counter() - t0
        log(f"[bench] autotuned {len(tuner.table)} size classes in {tune_s:.1f} s: lanes per "
            f"pixel {dict(sorted(collections.Counter(lpp).items()))}")
        items = list(zip(tess, params, bufs))
        for _ in range(a.warmup):
            b.run(items, lanes_per_pixel=lpp)
        walk, wall = [], []
        for _ in range(a.steps):
            for buf in bufs:
                buf.zero()
            w0 = time.perf_counter()
            _, tot = b.run(items, lanes_per_pixel=lpp)
            wall.append(time.perf_counter() - w0)
            walk.append(tot["t_walk_ms"])
    # the same maps one launch each
    w0 = time.perf_counter()
    for t, p, buf in zip(tess, params, bufs):
        t.map_device(p, buf, stats=False)
    for t in tess:
        t.sync()
    single_s = time.perf_counter() - w0
    for t in tess:
        t.close()
    for buf in bufs:
        buf.close()
    ms = 1e3 * float(np.mean(wall))
    out = {"metric": "sample-rays/s", "value": round(rays / (ms / 1e3), 1),
           "unit": "sample-rays/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32+f64" if prec == capi.PREC_MIXED else "f64",
           "data": "synthetic",
           "config": {"workload": f"{len(reqs)} halo cutouts (1e3-2e5 particles, 64-1024 px, "
                                  f"4-16 samples/pixel) in one batched launch",
                      "sample_rays": rays, "maps": len(reqs)},
           "lanes_per_pixel": {str(k): v for k, v in sorted(collections.Counter(lpp).items())},
           "autotune_s": round(tune_s, 2),
           "batched_walk_ms": round(float(np.mean(walk)), 3),
           "one_launch_per_map_s": round(single_s, 3),
           "batched_over_one_by_one": round(single_s / (ms / 1e3), 2)}
    print(json.dumps(out), flush=True)


def bench_serve(a):
    """Serving: a Poisson stream of cutout requests (particles, tets, map
    parameters) against the continuous-batching server; first all at once
    (capacity), then at 75% of that rate.  Reports p50/p99 latency from
    submission to the map on the host."""
    from sdtfe_amd import capi, serve
    prec = _precision(a)
    t0 = time.perf_counter()
    reqs = serve.make_requests(a.cutouts, seed=a.seed, workers=16)
    log(f"[bench] {len(reqs)} requests tessellated in {time.perf_counter() - t0:.1f} s")
    with capi.Server(hbm_budget_bytes=64 << 30, max_active=64, tile_budget=8192,
                     precision=prec) as srv:
        cap = serve.run_stream(srv, reqs, None, seed=a.seed)
    rate = 0.75 * cap["requests"] / cap["wall_s"]
    log(f"[bench] capacity {cap['requests'] / cap['wall_s']:.1f} maps/s; stream at {rate:.1f}/s")
    with capi.Server(hbm_budget_bytes=64 << 30, max_active=64, tile_budget=8192,
                     precision=prec) as srv:
        res = serve.run_stream(srv, reqs, rate, seed=a.seed + 1)
    out = {"metric": "sample-rays/s", "value": round(res["sample_rays_per_s"], 1),
           "unit": "sample-rays/s", "n_gpus": 1, "steps": 1, "warmup": 0,
           "ms_per_step": round(1e3 * res["wall_s"], 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None,
           "dtype": "f32+f64" if prec == capi.PREC_MIXED else "f64", "data": "synthetic",
           "config": {"workload": f"Poisson stream of {len(reqs)} cutout requests at "
                                  f"{rate:.1f}/s (75% of measured capacity)",
                      "requests": len(reqs)},
           "p50_ms": round(res["p50_ms"], 2), "p99_ms": round(res["p99_ms"], 2),
           "mean_ms": round(res["mean_ms"], 2), "queue_p99_ms": round(res["queue_p99_ms"], 2),
           "capacity_maps_per_s": round(cap["requests"] / cap["wall_s"], 2),
           "capacity_sample_rays_per_s": round(cap["sample_rays_per_s"], 1),
           "server": res["server"]}
    print(json.dumps(out), flush=True)


SERVE_RATE_HZ = 150.0  # arrivals per second of the serving leg in the default run


def serving_leg(n_requests, seed, rate_hz=SERVE_RATE_HZ):
    """Poisson stream of n_requests cutout requests at rate_hz through the
    continuous-batching server, once per precision: capacity (all requests
    at once), p50/p99 latency from submission to the map on the host, and
    the mixed maps' largest relative error against the FP64 maps."""
    from sdtfe_amd import capi, serve
    t0 = time.perf_counter()
    reqs = serve.make_requests(n_requests, seed=seed, workers=16)
    log(f"[bench] serving: {len(reqs)} requests tessellated in {time.perf_counter() - t0:.1f} s")
    rays = sum(r.rays for r in reqs)
    out = {"workload": f"Poisson stream of {len(reqs)} cutout requests (1e3-2e5 particles, "
                       f"64-1024 px, 4-16 samples/pixel) at {rate_hz:g}/s",
           "requests": len(reqs), "sample_rays": rays, "rate_hz": rate_hz}
    maps = {}
    for name, prec in (("fp64", capi.PREC_FP64), ("mixed", capi.PREC_MIXED)):
        kw = dict(hbm_budget_bytes=64 << 30, max_active=64, tile_budget=8192, precision=prec)
        with capi.Server(**kw) as srv:
            cap = serve.run_stream(srv, reqs, None, seed=seed)
        m = []
        with capi.Server(**kw) as srv:
            res = serve.run_stream(srv, reqs, rate_hz, seed=seed + 1, maps=m)
        maps[name] = m
        out[name] = {"p50_ms": round(res["p50_ms"], 2), "p99_ms": round(res["p99_ms"], 2),
                     "mean_ms": round(res["mean_ms"], 2),
                     "sample_rays_per_s": round(res["sample_rays_per_s"], 1),
                     "capacity_maps_per_s": round(cap["requests"] / cap["wall_s"], 1),
                     "capacity_sample_rays_per_s": round(cap["sample_rays_per_s"], 1)}
        log(f"[bench] serving {name}: p50 {res['p50_ms']:.1f} ms, p99 {res['p99_ms']:.1f} ms, "
            f"capacity {cap['requests'] / cap['wall_s']:.1f} maps/s")
    worst = 0.0
    for a_, b_ in zip(maps["mixed"], maps["fp64"]):
        e = rel_err(a_, b_)
        if e.size:
            worst = max(worst, float(e.max()))
    out["mixed"]["max_rel_err_vs_fp64"] = worst
    return out


def bench_volume(a):
    """Volumetric grid (compute_3d_density's interpolation, src/dtfe.c:
    667-735) of the 1,048,576 clustered particles' density: a 512^3 grid of
    the unit cube, 2 x 2 rays per column and 2 points per cell along each
    (1,048,576 rays, 268M field samples), FP64, accumulated in HBM.  The
    oracle computes the whole grid on the host cores, which is both the CPU
    baseline and a bit-for-bit check of the GPU grid."""
    from sdtfe_amd import capi
    from sdtfe_amd.tessellate import to_particle_data
    G, ss = a.vol_grid, a.vol_ss
    pos = particles("clustered")
    t0 = time.perf_counter()
    tetra = tessellation("clustered", pos, a.tess_cache)
    log(f"[bench] tessellated {pos.shape[0]} particles -> {tetra.shape[0]} tets in "
        f"{time.perf_counter() - t0:.1f} s")
    tess = capi.Tessellation(to_particle_data(pos), tetra)
    cells = G ** 3
    buf = capi.DeviceBuffer(cells * 8)
    p = tess.volume_params(G, ss, seed=a.seed)
    buf.zero()
    for _ in range(a.warmup):
        tess.volume_device(p, buf, stats=True)
    walk, st = [], None
    tess.sync()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        st = tess.volume_device(p, buf, stats=True)
        walk.append(st["t_walk_ms"])
    tess.sync()
    wall = time.perf_counter() - t0
    # one grid, untimed, into a clean buffer for the comparison
    buf.zero()
    tess.volume_device(p, buf, stats=False)
    tess.sync()
    grid = buf.to_host()
    buf.close()
    tess.close()
    rays = G * G * ss * ss
    ms = 1e3 * wall / a.steps
    kernel_ms = float(np.mean(walk))
    log(f"[bench] volume {G}^3 x{ss}: {ms:.2f} ms/grid (kernel {kernel_ms:.2f} ms), "
        f"{st['tet_steps'] / rays:.1f} tets/ray")
    # algorithmic bytes: 88 per tet crossed, plus one read and one write of
    # every cell each ray passes (the cells of its column inside the hull);
    # kernel time: HIP events around the launch sequence (with supersampling
    # > 1 the ray kernel and the fold kernel)
    cell_visits = rays * G
    alg = st["tet_steps"] * BYTES_PER_STEP + 16 * cell_visits
    achieved = alg / (kernel_ms / 1e3) / 1e9
    workload = (f"clustered-{pos.shape[0]} {G}^3 volumetric grid, {ss}x{ss} rays per column, "
                f"{ss} points per cell in z")
    kernel = ("volume_rays_kernel+volume_fold_kernel" if ss > 1 else "volume_fp64_kernel")
    tr = measured_traffic(kernel, workload)
    cpu = cpu_ref = None
    if not a.no_cpu:
        from oracle import oracle as O
        if not os.path.exists(O.LIB_PATH):
            O.build()
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        pd = to_particle_data(pos)
        log(f"[bench] CPU baseline: the whole grid on {threads} threads")
        t0 = time.perf_counter()
        want, st_o = O.volume(pd, tetra, G, 1.0, (0.5, 0.5, 0.5), ss, seed=a.seed, p_mass=1.0,
                              n_threads=threads)
        t_cpu = time.perf_counter() - t0
        same = bool(np.array_equal(grid, want))
        log(f"[bench] CPU {rays / t_cpu:.3e} rays/s ({t_cpu:.1f} s incl. density and "
            f"gradients); GPU grid bit-equal: {same}")
        cpu = {"value": rays / t_cpu, "unit": "rays/s", "cores": threads, "kind": "port",
               "sample": f"the whole {G}^3 grid (density, gradients, entry walks and rays), "
                         f"{t_cpu:.1f} s",
               "grid_bit_equal": same}
        if O.ref_available():
            # the reference's own code over the whole grid on one core
            rs = O.RefSetup(to_particle_data(pos), tetra)
            try:
                t0 = time.perf_counter()
                ref_grid, ref_rays = rs.volume(G, 1.0, (0.5, 0.5, 0.5), ss, a.seed)
                t_ref = time.perf_counter() - t0
            finally:
                rs.close()
            ref_same = bool(np.array_equal(grid, ref_grid))
            log(f"[bench] reference's own evaluate_3d_ray_on_grid: {ref_rays / t_ref:.3e} rays/s "
                f"on 1 core ({t_ref:.1f} s); GPU grid bit-equal: {ref_same}")
            cpu_ref = {"value": ref_rays / t_ref, "unit": "rays/s", "cores": 1,
                       "kind": "reference",
                       "sample": f"the whole {G}^3 grid, rays only ({t_ref:.1f} s)",
                       "grid_bit_equal": ref_same}
    out = {"metric": "rays/s", "value": round(rays / (ms / 1e3), 1), "unit": "rays/s",
           "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": workload, "grid": G, "supersampling": ss, "rays": rays,
                      "cells": cells},
           "cells_per_s": round(cells / (ms / 1e3), 1),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": None if tr is None else tr["bytes_per_launch"],
                        "kernel": kernel,
                        "bytes_per_step": BYTES_PER_STEP, "bytes_per_cell_visit": 16,
                        "tet_steps_per_launch": st["tet_steps"], "cell_visits": cell_visits,
                        "kernel_ms": round(kernel_ms, 3)},
           "cpu_baseline": cpu,
           "cpu_reference": cpu_ref,
           "stats": {k: st[k] for k in ("samples", "missed", "failed", "tet_steps",
                                        "perturb_steps")}}
    if tr is not None:
        out["roofline"]["physical_gbs"] = round(tr["bytes_per_launch"] / (kernel_ms / 1e3) / 1e9, 1)
    if cpu is not None:
        out["gpu_over_cpu"] = round(out["value"] / cpu["value"], 2)
    print(json.dumps(out), flush=True)


def roofline_entry(kernel, workload, f):
    """The walk kernel against HBM: physical bytes per launch (the committed
    FETCH_SIZE / WRITE_SIZE profile of this kernel and workload) over the
    kernel's time measured in this run with HIP events.  Beside it the
    algorithmic rate (88 B per tet crossed, which counts no reuse between
    rays and so exceeds the HBM peak when rays share records through the
    caches) and the unit the profile finds busiest."""
    algo = f["tet_steps"] * BYTES_PER_STEP / (f["kernel_ms"] / 1e3) / 1e9
    tr = measured_traffic(kernel, workload)
    phys = None if tr is None else tr["bytes_per_launch"] / (f["kernel_ms"] / 1e3) / 1e9
    out = dict(bound="hbm", basis="physical HBM bytes per launch / kernel time",
               achieved=None if phys is None else round(phys, 1), peak=HBM_PEAK_GBS,
               unit="GB/s", frac=None if phys is None else round(phys / HBM_PEAK_GBS, 4),
               traffic=None if tr is None else tr["bytes_per_launch"],
               kernel=kernel, kernel_ms=round(f["kernel_ms"], 3), walk_ms=round(f["walk_ms"], 3))
    out["algorithmic"] = dict(bytes_per_step=BYTES_PER_STEP, tet_steps_per_launch=f["tet_steps"],
                              gbs=round(algo, 1), over_hbm_peak=round(algo / HBM_PEAK_GBS, 3),
                              reuse=None if phys is None else round(algo / phys, 1))
    if tr is not None:
        out["traffic_profile"] = dict(instantiation=tr.get("kernel"), source=tr.get("source"),
                                      source_stamp=tr.get("source_stamp"), stale=tr["stale"],
                                      kernel_ms_rocprof=tr.get("kernel_ms_rocprof"))
        if tr.get("limiter"):
            out["limiter"] = tr["limiter"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="clustered",
                    choices=sorted(CONFIGS) + ["cutouts", "serve", "volume"])
    ap.add_argument("--vol-grid", type=int, default=512, help="volume: cells per side")
    ap.add_argument("--vol-ss", type=int, default=2, help="volume: supersampling")
    ap.add_argument("--cutouts", type=int, default=512, help="cutouts / requests (cutouts, serve)")
    ap.add_argument("--precision", default="auto", choices=["auto", "fp64", "mixed"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=256,
                    help="default run: requests of the serving leg (0: none)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--tess-cache", default=None, help="npz file caching the host tessellation")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (counter calibration)")
    a = ap.parse_args()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the path shards by "
              "whole maps or 4-row bands, which this library does not run across GPUs.",
              file=sys.stderr)
        sys.exit(2)

    if a.config in ("cutouts", "serve", "volume"):
        {"cutouts": bench_cutouts, "serve": bench_serve, "volume": bench_volume}[a.config](a)
        return

    from sdtfe_amd import capi
    from sdtfe_amd.tessellate import to_particle_data
    cfg = CONFIGS[a.config]
    G, S = cfg["grid"], cfg["samples"]
    log(f"[bench] generating {a.config} particles")
    pos = particles(a.config)
    t0 = time.perf_counter()
    tetra = tessellation(a.config, pos, a.tess_cache)
    log(f"[bench] tessellated {pos.shape[0]} particles -> {tetra.shape[0]} tets in "
        f"{time.perf_counter() - t0:.1f} s")
    assert capi.device_count() >= 1, "no HIP device"
    tess = capi.Tessellation(to_particle_data(pos), tetra)
    info = tess.info()
    log(f"[bench] resident: {info['device_bytes'] / 2**30:.2f} GiB, setup {info['t_setup_ms']:.0f} ms,"
        f" {info['n_tri']} hull faces, grid {info['grid_nx']}x{info['grid_ny']}")

    forms = {}
    want = ["fp64", "mixed"] if a.precision == "auto" else [a.precision]
    for name in want:
        prec = capi.PREC_MIXED if name == "mixed" else capi.PREC_FP64
        try:
            log(f"[bench] {name}: {a.warmup} warmup + {a.steps} timed maps")
            forms[name] = time_form(tess, cfg, prec, a.steps, a.warmup, a.seed)
            f = forms[name]
            log(f"[bench] {name}: {f['ms_per_step']:.1f} ms/map, walk {f['walk_ms']:.1f} ms, "
                f"{f['tet_steps'] / (G * G * S):.1f} tets/ray")
        except capi.SdtfeError as e:
            log(f"[bench] {name}: unavailable ({e})")
    tess.close()

    # the drop-in entry point end to end (host arrays in, host map out: upload,
    # density, gradients, records, entry grid, map, download; the tessellation
    # excluded), in the headline precision chosen below; reported, never the value
    dropin = {}
    for name in forms:
        os.environ["SDTFE_PRECISION"] = name
        ts = []
        try:
            for _ in range(2):  # the first call also warms the allocator pool
                pd_h = to_particle_data(pos)
                rho_h = np.zeros(G * G)
                capi.set_seed(a.seed)
                t0 = time.perf_counter()
                capi.compute_density(pd_h, tetra, G, cfg["box_len"], cfg["box_len_z"], rho_h, 1.0,
                                     cfg["center"], S, 0.0)
                ts.append(time.perf_counter() - t0)
        finally:
            os.environ.pop("SDTFE_PRECISION", None)
        dropin[name] = min(ts)
        log(f"[bench] {name}: drop-in compute_density end to end {1e3 * dropin[name]:.1f} ms")

    err = None
    if "fp64" in forms and "mixed" in forms:
        e = rel_err(forms["mixed"]["map"], forms["fp64"]["map"])
        err = dict(max=float(e.max()), p99=float(np.quantile(e, 0.99)),
                   median=float(np.median(e)))
        log(f"[bench] mixed vs fp64 per-pixel rel err: {err}")
    # headline form: the faster of FP64 (the reference's arithmetic) and the
    # mixed walk when that meets 1e-5 against the FP64 map
    ok = [k for k in forms if k == "fp64" or (err is not None and err["max"] <= 1e-5)]
    head = min(ok or list(forms), key=lambda k: forms[k]["ms_per_step"])
    f = forms[head]
    rays = G * G * S
    value = rays / (f["ms_per_step"] / 1e3)
    kernel = "map_mixed_kernel" if head == "mixed" else "map_fp64_kernel"
    workload = f"{a.config}-{pos.shape[0]} {G}x{G} map, {S} samples/pixel"
    roofline = roofline_entry(kernel, workload, f)

    serving = None
    if a.serve_requests > 0:
        serving = serving_leg(a.serve_requests, a.seed)

    cpu = None
    if not a.no_cpu:
        log(f"[bench] CPU baseline (oracle, ~{a.cpu_seconds:.0f} s)")
        cpu = cpu_baseline(pos, tetra, cfg, a.seed, a.cpu_seconds,
                           forms["fp64"]["map"] if "fp64" in forms else None)
        log(f"[bench] CPU {cpu['value']:.3e} sample-rays/s on {cpu['cores']} threads; "
            f"GPU/CPU {value / cpu['value']:.1f}x")
        if cpu["reference"]:
            r = cpu["reference"]
            log(f"[bench] reference's own walk {r['value']:.3e} sample-rays/s on 1 thread "
                f"({r['sample']}); equal to the port: {r['equal_to_port']}, to the GPU's FP64 "
                f"map: {r['equal_to_gpu_fp64']}")

    out = {
        "metric": "sample-rays/s", "value": round(value, 1), "unit": "sample-rays/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(f["ms_per_step"], 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": workload,
                   "particles": int(pos.shape[0]), "tets": int(tetra.shape[0]),
                   "grid": G, "samples_per_pixel": S, "sample_rays": rays,
                   "precision": head},
        "roofline": roofline,
        "cpu_baseline": None if cpu is None else {
            **{k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample")},
            "cpu_model": cpu_model()},
        "forms": {k: {"sample_rays_per_s": round(rays / (v["ms_per_step"] / 1e3), 1),
                      "ms_per_step": round(v["ms_per_step"], 3),
                      "walk_ms": round(v["walk_ms"], 3),
                      "kernel_ms": round(v["kernel_ms"], 3),
                      "tet_steps": v["tet_steps"],
                      "missed": v["stats"]["missed"], "failed": v["stats"]["failed"],
                      "perturb_steps": v["stats"]["perturb_steps"],
                      "fp64_redo": v["stats"]["fp64_redo"]} for k, v in forms.items()},
        "mixed_vs_fp64_rel_err": err,
        "pcie_inclusive": {k: {"ms": round(1e3 * v, 2),
                               "sample_rays_per_s": round(rays / v, 1)}
                           for k, v in dropin.items()},
        "serving": serving,
    }
    if cpu is not None:
        out["gpu_over_cpu"] = round(value / cpu["value"], 2)
        out["fp64_vs_cpu_max_rel_err"] = cpu["gpu_vs_cpu_max_rel_err"]
        out["cpu_reference"] = cpu["reference"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
