"""Synthetic particle sets for the surface-density workloads.

All sets are float64 positions drawn from numpy's PCG64 ``default_rng`` with
the stated seed; particles have unit mass.

* uniform          -- N points uniform in the unit box (the 32,768-particle,
                      seed-0 map: 256x256 pixels, 4 samples per pixel).
* clustered        -- 30% uniform background + 70% in 200 halos with NFW-like
                      radial profiles (concentration U[5, 10], log-uniform
                      masses in [1e-4, 1e-2] of the total renormalised to 70%,
                      virial radius 0.05 * (M / 1e-2)^(1/3)), wrapped into the
                      unit box (1,048,576 particles: 2048x2048 pixels, 16
                      samples per pixel).
* jittered_lattice -- nx*ny*nz lattice of spacing h with N(0, (0.3 h)^2)
                      displacements, 5% of the points re-drawn into halos.
* cutout           -- one halo cutout: NFW-like halo centred in the unit box
                      plus a uniform background.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


def _nfw_radii(rng: np.random.Generator, n: int, c: float) -> np.ndarray:
    """Radii in units of R_vir with the truncated NFW enclosed-mass CDF
    m(x) = ln(1 + c x) - c x / (1 + c x), x <= 1, inverted on a table."""
    x = np.linspace(0.0, 1.0, 4097)
    m = np.log1p(c * x) - c * x / (1.0 + c * x)
    m /= m[-1]
    return np.interp(rng.random(n), m, x)


def _directions(rng: np.random.Generator, n: int) -> np.ndarray:
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v


def _halos(rng: np.random.Generator, n_points: int, n_halos: int, box: np.ndarray,
           r_ref: float = 0.05, m_ref: float = 1e-2) -> np.ndarray:
    """n_points particles spread over n_halos NFW-like halos in ``box``."""
    centres = rng.random((n_halos, 3)) * box
    conc = rng.uniform(5.0, 10.0, n_halos)
    masses = np.exp(rng.uniform(math.log(1e-4), math.log(1e-2), n_halos))
    share = masses / masses.sum()
    counts = np.floor(share * n_points).astype(np.int64)
    # hand the rounding remainder to the largest fractional parts
    rem = n_points - int(counts.sum())
    if rem > 0:
        order = np.argsort(-(share * n_points - counts), kind="stable")
        counts[order[:rem]] += 1
    rvir = r_ref * np.cbrt(masses / m_ref)
    out = np.empty((n_points, 3))
    k = 0
    for h in range(n_halos):
        n = int(counts[h])
        if n == 0:
            continue
        r = _nfw_radii(rng, n, float(conc[h])) * rvir[h]
        out[k:k + n] = centres[h] + _directions(rng, n) * r[:, None]
        k += n
    return np.mod(out, box)


def uniform(n: int = 32768, seed: int = 0) -> np.ndarray:
    return np.random.default_rng(seed).random((n, 3))


def clustered(n: int = 1_048_576, seed: int = 0, n_halos: int = 200,
              background: float = 0.3) -> np.ndarray:
    rng = np.random.default_rng(seed)
    n_bg = int(round(background * n))
    bg = rng.random((n_bg, 3))
    halo = _halos(rng, n - n_bg, n_halos, np.ones(3))
    pos = np.concatenate([bg, halo])
    return pos[rng.permutation(n)]


def jittered_lattice(nx: int = 256, ny: int = 256, nz: int = 128, seed: int = 0,
                     sigma: float = 0.3, halo_fraction: float = 0.05,
                     n_halos: int = 200) -> np.ndarray:
    rng = np.random.default_rng(seed)
    h = 1.0 / nx
    box = np.array([nx * h, ny * h, nz * h])
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"),
                 axis=-1).reshape(-1, 3).astype(np.float64)
    pos = (g + 0.5) * h + rng.normal(scale=sigma * h, size=g.shape)
    pos = np.mod(pos, box)
    n = pos.shape[0]
    n_h = int(round(halo_fraction * n))
    idx = rng.choice(n, n_h, replace=False)
    pos[idx] = _halos(rng, n_h, n_halos, box)
    return pos


def cutout(n: int, seed: int, background: float = 0.2, r_vir: float = 0.45,
           conc: float | None = None) -> np.ndarray:
    """One halo cutout in the unit box: NFW-like halo of radius r_vir about
    (0.5, 0.5, 0.5) plus a uniform background fraction."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(5.0, 10.0) if conc is None else conc
    n_bg = max(8, int(round(background * n)))
    n_h = n - n_bg
    r = _nfw_radii(rng, n_h, c) * r_vir
    halo = 0.5 + _directions(rng, n_h) * r[:, None]
    return np.concatenate([rng.random((n_bg, 3)), halo])


@dataclass
class MapSpec:
    """Arguments of one compute_density call besides the particles."""
    grid_dim: int
    n_samples: int
    box_len: float = 1.0
    box_len_z: float = 1.0
    center: tuple = (0.5, 0.5, 0.5)
    delta_sample: float = 0.0
    p_mass: float = 1.0
    seed: int = 0


@dataclass
class CutoutSpec:
    n_particles: int
    seed: int
    map: MapSpec = field(default_factory=lambda: MapSpec(64, 4))


def batch_specs(n_cutouts: int = 512, seed: int = 0, n_lo: float = 1e3, n_hi: float = 2e5,
                g_lo: float = 64, g_hi: float = 1024, s_lo: int = 4, s_hi: int = 16):
    """The batched-cutout workload: particle counts log-uniform in
    [n_lo, n_hi], grids log-uniform in [g_lo, g_hi] rounded to a multiple of
    4, samples per pixel uniform in {s_lo..s_hi}; cutout i uses seed i."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_cutouts):
        n = int(round(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi)))))
        g = int(round(math.exp(rng.uniform(math.log(g_lo), math.log(g_hi))) / 4.0)) * 4
        s = int(rng.integers(s_lo, s_hi + 1))
        out.append(CutoutSpec(n, i, MapSpec(max(g, 4), s, seed=i)))
    return out


C1 = dict(name="uniform-32768", n=32768, spec=MapSpec(256, 4))
C2 = dict(name="clustered-1048576", n=1_048_576, spec=MapSpec(2048, 16))
C3 = dict(name="lattice-8388608", n=8_388_608,
          spec=MapSpec(4096, 8, box_len=1.0, box_len_z=0.5, center=(0.5, 0.5, 0.25)))
