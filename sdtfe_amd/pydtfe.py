"""Drop-in for the reference's ``pydtfe`` module: ``compute_surface_density``,
``compute_3d_density``, ``interpolate_field2grid`` and
``compute_particle_density`` (python/pydtfe.cpp:16-203).

Same function name, argument names, defaults and errors as the pybind11
binding (python/pydtfe.cpp:16-66, 207-261); the map itself is computed by
``compute_density`` of libsdtfe_amd.so on the GPU (the same C entry point the
reference binding calls, src/dtfe.h:13-15).

Pre-processing restated from the binding:

* positions forcecast to float32, Fortran order (python/pydtfe.cpp:17), then
  widened to the N x 4 float64 buffer (convert_to_qhdata, src/utils.c:8-16);
* shape check -> ValueError("incorrect shape of particle position array")
  (python/pydtfe.cpp:32-34);
* n_shuffle = size_t(N * float(sample_factor)); fewer than 5 ->
  ValueError("not enough particles") (python/pydtfe.cpp:39-42);
* partial Fisher-Yates shuffle of the first n_shuffle slots
  (shuffle_buff, src/utils.c:18-30) -- always applied, as the reference does;
* particle_mass /= sample_factor, narrowed to float at the C call
  (python/pydtfe.cpp:45, 59);
* optional rotation about ``center`` (rotate3d, src/utils.c:33-79);
* Qhull Delaunay of the first n_shuffle particles ("Qz Qt Qbb").

Randomness: the reference draws the shuffle and the Monte-Carlo samples from
the unseeded libc rand().  Here both come from the counter-based stream with
``seed`` (keyword-only extension; default 0), so results are reproducible.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import capi
from .tessellate import triangulate

__all__ = ["compute_surface_density", "compute_3d_density", "interpolate_field2grid",
           "compute_particle_density"]

_SHUFFLE_PIXEL = (1 << 63) + 0x5348  # stream key reserved for the shuffle


def convert_to_qhdata(pos_f32: np.ndarray) -> np.ndarray:
    """(N, 3) float32 -> N x 4 float64 (x, y, z, 0) (src/utils.c:8-16)."""
    out = np.zeros((pos_f32.shape[0], 4), dtype=np.float64)
    out[:, :3] = pos_f32.astype(np.float64)
    return out


def _draws(seed: int, n: int) -> np.ndarray:
    """n 31-bit draws of the shuffle stream (same SplitMix64 stream as the
    GPU sample stream, under a reserved pixel key)."""
    M = (1 << 64) - 1
    g = 0x9E3779B97F4A7C15

    def mix(z):
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9))
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB))
        return z ^ (z >> np.uint64(31))

    with np.errstate(over="ignore"):
        k = int(mix(np.uint64((seed + g) & M)))
        k = int(mix(np.uint64(k ^ _SHUFFLE_PIXEL)))
        k = int(mix(np.uint64(k)))  # sample 0
        d = np.arange(1, n + 1, dtype=np.uint64) * np.uint64(g) + np.uint64(k)
        return (mix(d) >> np.uint64(33)).astype(np.int64)


def shuffle_buff(data: np.ndarray, n_in_buff: int, n_shuffle: int, seed: int) -> None:
    """In place: for i = 1..n_shuffle-1 swap row i-1 with row
    r = draw % (n_in_buff - i) + i (src/utils.c:18-30)."""
    if n_shuffle <= 1:
        return
    r = _draws(seed, n_shuffle - 1)
    for i in range(1, n_shuffle):
        j = int(r[i - 1] % (n_in_buff - i)) + i
        if j != i - 1:
            tmp = data[j].copy()
            data[j] = data[i - 1]
            data[i - 1] = tmp


def rotation_matrix(theta) -> np.ndarray:
    """Qt2 = (Qx Qy) Qz with each product summed m = 0..2 from 0
    (src/utils.c:43-69)."""
    cx, sx = math.cos(theta[0]), math.sin(theta[0])
    cy, sy = math.cos(theta[1]), math.sin(theta[1])
    cz, sz = math.cos(theta[2]), math.sin(theta[2])
    qx = [[1.0, 0.0, 0.0], [0.0, cx, -sx], [0.0, sx, cx]]
    qy = [[cy, 0.0, sy], [0.0, 1.0, 0.0], [-sy, 0.0, cy]]
    qz = [[cz, -sz, 0.0], [sz, cz, 0.0], [0.0, 0.0, 1.0]]

    def mul(a, b):
        out = [[0.0] * 3 for _ in range(3)]
        for i in range(3):
            for j in range(3):
                acc = 0.0
                for m in range(3):
                    acc = acc + a[i][m] * b[m][j]
                out[i][j] = acc
        return out

    return np.array(mul(mul(qx, qy), qz))


def rotate3d(data: np.ndarray, n: int, theta, center) -> None:
    """Rotate the first n rows about ``center`` in place (src/utils.c:71-78)."""
    q = rotation_matrix(theta)
    c = [float(v) for v in center]
    x = data[:n, 0] - c[0]
    y = data[:n, 1] - c[1]
    z = data[:n, 2] - c[2]
    nx = x * q[0, 0] + y * q[0, 1] + z * q[0, 2] + c[0]
    ny = x * q[1, 0] + y * q[1, 1] + z * q[1, 2] + c[1]
    nz = x * q[2, 0] + y * q[2, 1] + z * q[2, 2] + c[2]
    data[:n, 0], data[:n, 1], data[:n, 2] = nx, ny, nz


def compute_surface_density(particle_pos, center, grid_dim, sample_factor, box_width, box_depth,
                            particle_mass, mc_box_width, mc_sample_count, rotation_angle=None,
                            *, seed: int = 0):
    """Surface density map (grid_dim x grid_dim, float64, mass/length^2) of
    the DTFE field of ``particle_pos`` integrated over
    [center_z - box_depth/2, center_z + box_depth/2]; see the reference's
    docstring (python/pydtfe.cpp:207-250)."""
    pos = np.asarray(particle_pos, dtype=np.float32, order="F")
    if pos.ndim != 2 or pos.shape[1] != 3:
        raise ValueError("incorrect shape of particle position array")
    n = pos.shape[0]
    pd = convert_to_qhdata(pos)
    n_shuffle = int(np.float32(n) * np.float32(sample_factor))
    if n_shuffle < 5:
        raise ValueError("not enough particles")
    shuffle_buff(pd, n, n_shuffle, seed)
    mass = float(particle_mass) / float(np.float32(sample_factor))
    if rotation_angle is not None:
        rotate3d(pd, n_shuffle, rotation_angle, center)
    pd = np.ascontiguousarray(pd[:n_shuffle])
    tetra = triangulate(pd, gpu=True)
    g = int(grid_dim)
    rho = np.zeros(g * g)
    capi.set_seed(seed)
    capi.compute_density(pd, tetra, g, float(box_width), float(box_depth), rho,
                         float(np.float32(mass)), [float(c) for c in center],
                         int(mc_sample_count), float(mc_box_width))
    return rho.reshape(g, g)


def _positions(particle_pos) -> np.ndarray:
    pos = np.asarray(particle_pos, dtype=np.float32, order="F")
    if pos.ndim != 2 or pos.shape[1] != 3:
        raise ValueError("incorrect shape of particle position array")
    return pos


def compute_3d_density(particle_pos, center, grid_dim, sample_factor, box_length, particle_mass,
                       supersampling=1, *, seed: int = 0):
    """Volumetric density (grid_dim^3, float64, mass/length^3, indexed
    [x, y, z]) of the DTFE field of ``particle_pos`` on the cube of side
    ``box_length`` about ``center``; each cell averages supersampling^3
    points.  Restates python/pydtfe.cpp:68-111: float32 positions, shuffle
    of the first N * sample_factor particles, mass / sample_factor, Qhull,
    then compute_3d_density (src/dtfe.h:20-21) of libsdtfe_amd.so."""
    pos = _positions(particle_pos)
    n = pos.shape[0]
    pd = convert_to_qhdata(pos)
    n_shuffle = int(np.float32(n) * np.float32(sample_factor))
    if n_shuffle < 5:
        raise ValueError("not enough particles")
    shuffle_buff(pd, n, n_shuffle, seed)
    mass = float(particle_mass) / float(np.float32(sample_factor))
    pd = np.ascontiguousarray(pd[:n_shuffle])
    tetra = triangulate(pd, gpu=True)
    g = int(grid_dim)
    rho = np.zeros(g ** 3)
    capi.set_seed(seed)
    capi.compute_3d_density(pd, tetra, g, float(box_length), rho, float(np.float32(mass)),
                            [float(c) for c in center], int(supersampling))
    return rho.reshape(g, g, g)


def interpolate_field2grid(particle_pos, particle_value, center, grid_dim, sample_factor,
                           box_length, supersampling=1, *, seed: int = 0):
    """A particle field (float32 values, one per particle) sampled on the
    grid_dim^3 grid, as compute_3d_density samples the density.  Restates
    python/pydtfe.cpp:113-163 (the same errors, sample_factor 1 only), then
    interpolate_to_3d_grid (src/dtfe.h:17-18) of libsdtfe_amd.so."""
    pos = _positions(particle_pos)
    n = pos.shape[0]
    val = np.asarray(particle_value, dtype=np.float32)
    if n != val.shape[0]:
        raise ValueError("different length of particle position and value arrays")
    pd = convert_to_qhdata(pos)
    if np.float32(sample_factor) != np.float32(1.0):
        raise ValueError("`sample_factor != 1.0f` not supported at the moment")
    tetra = triangulate(pd, gpu=True)
    pd[:, 3] = val.astype(np.float64)
    g = int(grid_dim)
    grid = np.zeros(g ** 3)
    capi.set_seed(seed)
    capi.interpolate_to_3d_grid(pd, tetra, g, float(box_length), grid,
                                [float(c) for c in center], int(supersampling))
    return grid.reshape(g, g, g)


def compute_particle_density(particle_pos, particle_mass):
    """DTFE density at each particle (float32, mass/length^3): Qhull, then
    pre_compute_vol (src/triangulate.h:26) of libsdtfe_amd.so; restates
    python/pydtfe.cpp:165-191."""
    pos = _positions(particle_pos)
    pd = convert_to_qhdata(pos)
    tetra = triangulate(pd, gpu=True)
    ptr = C.c_void_p(pd.ctypes.data)
    capi.lib().pre_compute_vol(C.byref(ptr), pd.shape[0], tetra, tetra.shape[0],
                               float(np.float32(particle_mass)))
    capi.check_dropin("pre_compute_vol")
    return pd[:, 3].astype(np.float32)
