"""Autotuner for the batched launch: picks the sample tiling (lanes per
pixel, hence rays in flight per wavefront: 64 lanes x S/L samples each) per
size class of map, by timing candidate launches on a representative item.

Size class = (grid bucket, samples per pixel, tets-per-column bucket); the
choice is cached in a JSON file so later batches reuse it.
"""
from __future__ import annotations

import json
import math
import os

from . import capi

CANDIDATES = (1, 2, 4, 8, 16)


def size_class(grid: int, samples: int, n_tetra: int) -> str:
    g = int(round(math.log2(max(grid, 1))))
    c = int(round(math.log2(max(n_tetra, 1)) / 2))  # ~ log4 of the tets
    return f"g{g}-s{samples}-t{c}"


class Autotuner:
    def __init__(self, cache_path: str | None = None, candidates=CANDIDATES, repeats: int = 2):
        self.cache_path = cache_path
        self.candidates = tuple(candidates)
        self.repeats = repeats
        self.table: dict[str, dict] = {}
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as f:
                self.table = json.load(f)

    def save(self):
        if self.cache_path:
            with open(self.cache_path, "w") as f:
                json.dump(self.table, f, indent=1, sort_keys=True)

    def lanes_for(self, batch: capi.Batch, tess: capi.Tessellation, params: capi.MapParams):
        key = size_class(params.grid_dim, params.n_mc_samp, tess.nt)
        if key not in self.table:
            self.table[key] = self.tune(batch, tess, params)
        return int(self.table[key]["lanes_per_pixel"])

    def tune(self, batch, tess, params) -> dict:
        npix = params.grid_dim * params.grid_dim
        buf = capi.DeviceBuffer(npix * 8)
        timings = {}
        try:
            for lpp in self.candidates:
                best = math.inf
                for _ in range(self.repeats):
                    buf.zero()
                    _, tot = batch.run([(tess, params, buf)], lanes_per_pixel=[lpp])
                    best = min(best, tot["t_walk_ms"])
                timings[lpp] = best
        finally:
            buf.close()
        pick = min(timings, key=timings.get)
        # JSON keys are strings
        return {"lanes_per_pixel": pick, "walk_ms": {str(k): v for k, v in timings.items()}}
