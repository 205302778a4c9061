"""ctypes access to the CPU oracle (liboracle.so) and the reference build
(_ref/libdtfe_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libdtfe_ref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


class Stats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "samples", "missed", "failed", "no_intersection", "perturb_exhausted",
        "no_isect_perturb", "perturb_steps", "tet_steps")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


def build(ref: bool = False) -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.sdo_draw31.restype = C.c_uint32
        L.sdo_draw31.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32]
        L.sdo_pre_compute_vol.argtypes = [_dp, C.c_int, _ip, C.c_int, C.c_float]
        L.sdo_tet_volumes.argtypes = [_dp, _ip, C.c_int, _dp]
        L.sdo_gradients.argtypes = [_ip, C.c_int, _dp, _dp]
        L.sdo_forward_faces.restype = C.c_int
        L.sdo_forward_faces.argtypes = [_ip, C.c_int, _dp, _up]
        L.sdo_copy_fh_triangles.argtypes = [_ip, C.c_int, _up, C.c_int, _ip]
        L.sdo_pt_loc_2d.restype = C.c_int
        L.sdo_pt_loc_2d.argtypes = [_ip, _dp, C.c_double, C.c_double, C.c_int]
        L.sdo_plucker.restype = C.c_int
        L.sdo_plucker.argtypes = [C.c_double, C.c_double, C.c_double,
                                  C.POINTER(C.c_void_p * 4), C.POINTER(C.c_int),
                                  C.POINTER(C.c_int), C.POINTER(C.c_double),
                                  C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.sdo_shoot_ray.restype = C.c_double
        L.sdo_shoot_ray.argtypes = [C.c_int, _ip, _dp, _dp, C.c_double, C.c_double,
                                    C.c_double, C.c_double, C.c_double, C.c_uint64,
                                    C.c_uint64, C.c_uint32, C.POINTER(C.c_int),
                                    C.POINTER(Stats)]
        L.sdo_compute_density.argtypes = [_dp, C.c_int, _ip, C.c_int, C.c_int, C.c_double,
                                          C.c_double, _dp, C.c_float, C.c_double,
                                          C.c_double, C.c_double, C.c_int, C.c_double,
                                          C.c_uint64, C.c_int, C.POINTER(Stats), C.c_void_p]
        L.sdo_map_bands.restype = C.c_int64
        L.sdo_map_bands.argtypes = [_dp, _ip, _dp, _ip, C.c_int, C.c_double, C.c_double, _dp,
                                    C.c_double, C.c_double, C.c_double, C.c_int, C.c_double,
                                    C.c_uint64, C.c_int, C.c_int, C.c_int, C.POINTER(Stats),
                                    C.c_void_p]
        L.sdo_ray_3d.restype = C.c_int
        L.sdo_ray_3d.argtypes = [C.c_int, _ip, _dp, _dp, C.c_double, C.c_double, C.c_double,
                                 C.c_double, C.c_int, C.c_double, C.c_uint, C.c_uint64,
                                 C.c_uint64, C.c_uint32, _dp, C.POINTER(Stats)]
        L.sdo_volume_columns.argtypes = [_dp, _ip, _dp, _ip, C.c_int, C.c_double, _dp,
                                         C.c_double, C.c_double, C.c_double, C.c_uint,
                                         C.c_uint64, C.c_void_p, C.c_int, C.POINTER(Stats)]
        L.sdo_interpolate_3d.argtypes = [_dp, C.c_int, _ip, C.c_int, C.c_int, C.c_double, _dp,
                                         C.c_double, C.c_double, C.c_double, C.c_uint,
                                         C.c_uint64, C.c_int, C.POINTER(Stats)]
        L.sdo_compute_3d_density.argtypes = [_dp, C.c_int, _ip, C.c_int, C.c_int, C.c_double,
                                             _dp, C.c_float, C.c_double, C.c_double,
                                             C.c_double, C.c_uint, C.c_uint64, C.c_int,
                                             C.POINTER(Stats)]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    """The reference C build (single-threaded) with the stream interposer."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError("reference build missing (make -C oracle ref)")
        R = C.CDLL(REF_PATH)
        R.refh_set_map.argtypes = [C.c_uint64, C.c_int, C.c_int]
        R.refh_set_ray.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32]
        R.refh_sample_draws.restype = C.c_int64
        R.refh_cd_size.restype = C.c_size_t
        R.compute_density.argtypes = [_dp, C.c_int, _ip, C.c_int, C.c_int, C.c_double,
                                      C.c_double, _dp, C.c_float, C.c_double, C.c_double,
                                      C.c_double, C.c_int, C.c_double]
        R.pre_compute_vol.argtypes = [C.POINTER(C.c_void_p), C.c_int, _ip, C.c_int, C.c_float]
        R.gradients.restype = C.c_void_p
        R.gradients.argtypes = [_ip, C.c_int, _dp]
        R.create_bitmap.restype = C.c_void_p
        R.create_bitmap.argtypes = [C.c_int]
        R.find_forward_hull.restype = C.c_int
        R.find_forward_hull.argtypes = [_ip, C.c_int, _dp, C.c_int, C.c_void_p]
        R.copy_fh_triangles.restype = C.c_void_p
        R.copy_fh_triangles.argtypes = [_ip, C.c_int, _dp, C.c_int, C.c_int, C.c_void_p]
        R.pt_loc_2d.restype = C.c_int
        R.pt_loc_2d.argtypes = [C.c_void_p, _dp, C.POINTER(C.c_int), C.c_double, C.c_double]
        R.shoot_3d_ray.restype = C.c_double
        R.shoot_3d_ray.argtypes = [C.c_int, _ip, C.c_void_p, _dp, C.c_double, C.c_double,
                                   C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                   C.POINTER(C.c_int)]
        R.rayTetraPlucker_full.restype = C.c_int
        R.rayTetraPlucker_full.argtypes = [_dp, C.POINTER(C.c_void_p * 4), C.POINTER(C.c_int),
                                           C.POINTER(C.c_int), _dp, _dp, C.POINTER(C.c_int)]
        R.refh_rand_calls.restype = C.c_int64
        R.refh_volume.restype = C.c_int64
        R.refh_volume.argtypes = [_dp, _ip, C.c_void_p, C.c_void_p, C.c_int, C.c_double, _dp,
                                  C.c_double, C.c_double, C.c_double, C.c_uint]
        R.refh_map_bands.restype = C.c_int64
        R.refh_map_bands.argtypes = [_dp, _ip, C.c_void_p, C.c_void_p, C.c_int, C.c_double,
                                     C.c_double, _dp, C.c_double, C.c_double, C.c_double,
                                     C.c_int, C.c_double, C.c_uint64, C.c_int, C.c_int]
        R.interpolate_to_3d_grid.argtypes = [_dp, C.c_int, _ip, C.c_int, C.c_int, C.c_double,
                                             _dp, C.c_double, C.c_double, C.c_double, C.c_uint]
        R.compute_3d_density.argtypes = [_dp, C.c_int, _ip, C.c_int, C.c_int, C.c_double, _dp,
                                         C.c_float, C.c_double, C.c_double, C.c_double,
                                         C.c_uint]
        R.evaluate_3d_ray_on_grid.argtypes = [C.c_int, _ip, C.c_void_p, _dp, C.c_double,
                                              C.c_double, C.c_double, C.c_double, C.c_int,
                                              C.c_double, C.POINTER(C.c_int), C.c_uint,
                                              C.c_void_p]
        R.rotate3d.argtypes = [_dp, C.c_size_t, _dp, _dp]
        R.convert_to_qhdata.restype = C.c_void_p
        R.convert_to_qhdata.argtypes = [np.ctypeslib.ndpointer(dtype=np.float32,
                                                               flags="C_CONTIGUOUS"), C.c_size_t]
        R.free = C.CDLL(None).free
        R.free.argtypes = [C.c_void_p]
        _ref = R
    return _ref


# ------------------------------------------------------------- oracle calls

def draw31(seed, pixel, sample, draw):
    return int(lib().sdo_draw31(seed, pixel, sample, draw))


def density(pd: np.ndarray, tetra: np.ndarray, p_mass: float) -> np.ndarray:
    """rho per vertex (writes pd[:, 3] in place, like the reference)."""
    lib().sdo_pre_compute_vol(pd, pd.shape[0], tetra, tetra.shape[0], p_mass)
    return pd[:, 3].copy()


def tet_volumes(pd, tetra):
    out = np.empty(tetra.shape[0])
    lib().sdo_tet_volumes(pd, tetra, tetra.shape[0], out)
    return out


def gradients(tetra, pd):
    out = np.empty((tetra.shape[0], 3))
    lib().sdo_gradients(tetra, tetra.shape[0], pd, out)
    return out


def forward_faces(tetra, pd):
    mask = np.zeros(tetra.shape[0], dtype=np.uint8)
    n = lib().sdo_forward_faces(tetra, tetra.shape[0], pd, mask)
    return n, mask


def triangles(tetra, pd):
    n, mask = forward_faces(tetra, pd)
    tri = np.zeros((max(n, 1), 8), dtype=np.int32)
    lib().sdo_copy_fh_triangles(tetra, tetra.shape[0], mask, n, tri)
    return tri[:n]


def pt_loc_2d(tri, pd, qx, qy, max_steps=1 << 24):
    return int(lib().sdo_pt_loc_2d(tri, pd, qx, qy, max_steps))


def shoot_ray(entry, tetra, nabla, pd, x, y, off_z, lz, dxy, seed, pixel, sample,
              stats=None):
    crossed = C.c_int(0)
    v = lib().sdo_shoot_ray(entry, tetra, nabla, pd, x, y, off_z, lz, dxy, seed, pixel,
                            sample, C.byref(crossed),
                            C.byref(stats) if stats is not None else None)
    return float(v), int(crossed.value)


def compute_density(pd, tetra, grid_dim, box_len, box_len_z, p_mass, center, n_samp,
                    delta_sample=0.0, seed=0, n_threads=0, rho=None, want_navg=False):
    """Oracle compute_density: returns (rho, stats[, n_avg]); pd[:,3] <- rho."""
    if rho is None:
        rho = np.zeros(grid_dim * grid_dim)
    st = Stats()
    navg = np.zeros(grid_dim * grid_dim, dtype=np.int32) if want_navg else None
    lib().sdo_compute_density(pd, pd.shape[0], tetra, tetra.shape[0], grid_dim, box_len,
                              box_len_z, rho, p_mass, center[0], center[1], center[2],
                              n_samp, delta_sample, seed, n_threads, C.byref(st),
                              navg.ctypes.data if navg is not None else None)
    if want_navg:
        return rho, st.as_dict(), navg
    return rho, st.as_dict()


def map_bands(pd, tetra, nabla, tri, grid_dim, box_len, box_len_z, center, n_samp,
              delta_sample=0.0, seed=0, band_begin=0, band_step=1, n_threads=0, rho=None,
              n_avg=None):
    """Pixels of every band_step-th 4-row band from band_begin; n_avg (int32,
    grid_dim^2, optional) receives those pixels' sample counts."""
    if rho is None:
        rho = np.zeros(grid_dim * grid_dim)
    st = Stats()
    n = lib().sdo_map_bands(pd, tetra, nabla, tri, grid_dim, box_len, box_len_z, rho,
                            center[0], center[1], center[2], n_samp, delta_sample, seed,
                            band_begin, band_step, n_threads, C.byref(st),
                            n_avg.ctypes.data if n_avg is not None else None)
    return rho, int(n), st.as_dict()


def band_rows(grid_dim, band_begin, band_step):
    """Boolean mask (grid_dim^2) of the pixels map_bands computes."""
    rows = np.zeros(grid_dim, dtype=bool)
    for b in range(band_begin, grid_dim // 4, band_step):
        rows[4 * b:4 * b + 4] = True
    return np.repeat(rows, grid_dim)


def volume(pd, tetra, grid_dim, box_len, center, supersampling, seed=0, p_mass=None,
           grid=None, n_threads=0):
    """Oracle volumetric grid (grid_dim^3, [x][y][z]): compute_3d_density when
    p_mass is given (pd[:, 3] <- rho), else interpolate_to_3d_grid of the
    field in pd[:, 3].  Returns (grid, stats)."""
    if grid is None:
        grid = np.zeros(grid_dim ** 3)
    st = Stats()
    args = (grid_dim, box_len, grid)
    if p_mass is None:
        lib().sdo_interpolate_3d(pd, pd.shape[0], tetra, tetra.shape[0], *args, center[0],
                                 center[1], center[2], supersampling, seed, n_threads,
                                 C.byref(st))
    else:
        lib().sdo_compute_3d_density(pd, pd.shape[0], tetra, tetra.shape[0], *args, p_mass,
                                     center[0], center[1], center[2], supersampling, seed,
                                     n_threads, C.byref(st))
    return grid, st.as_dict()


def volume_columns(pd, tetra, nabla, tri, grid_dim, box_len, center, supersampling, seed=0,
                   entries=None, grid=None, n_threads=0):
    """The column loop alone, given the field, gradients and triangles;
    entries (optional int32, one per ray in (column, is, js) order) replace
    the entry walk."""
    if grid is None:
        grid = np.zeros(grid_dim ** 3)
    st = Stats()
    e = None
    if entries is not None:
        entries = np.ascontiguousarray(entries, dtype=np.int32)
        e = entries.ctypes.data
    lib().sdo_volume_columns(pd, tetra, nabla, tri, grid_dim, box_len, grid, center[0],
                             center[1], center[2], supersampling, seed, e, n_threads,
                             C.byref(st))
    return grid, st.as_dict()


def volume_ray_xy(grid_dim, box_len, center, supersampling):
    """Ray positions of the volumetric path in (column, is, js) order
    (src/dtfe.c:688-713), computed with the reference's operation order."""
    G, ss = grid_dim, supersampling
    dxy = box_len / G
    off_x = -(box_len / 2.0) + center[0]
    off_y = -(box_len / 2.0) + center[1]
    ds = dxy / ss
    i = np.repeat(np.arange(G), G * ss * ss).astype(np.float64)
    j = np.tile(np.repeat(np.arange(G), ss * ss), G).astype(np.float64)
    is_ = np.tile(np.repeat(np.arange(ss), ss), G * G).astype(np.float64)
    js = np.tile(np.arange(ss), G * G * ss).astype(np.float64)
    x = i * dxy + off_x + is_ * ds + ds / 2.
    y = j * dxy + off_y + js * ds + ds / 2.
    return x, y


class RefSetup:
    """The reference's own field and triangle list of a tessellation
    (pre_compute_vol, gradients, find_forward_hull, copy_fh_triangles from
    the compiled reference), kept as C pointers for ref_map_bands."""

    def __init__(self, pd, tetra, p_mass=1.0):
        R = ref()
        self.R = R
        self.pd = np.ascontiguousarray(pd, dtype=np.float64)
        self.tetra = np.ascontiguousarray(tetra, dtype=np.int32)
        n, nt = self.pd.shape[0], self.tetra.shape[0]
        ptr = C.c_void_p(self.pd.ctypes.data)
        R.pre_compute_vol(C.byref(ptr), n, self.tetra, nt, p_mass)
        self.nabla = R.gradients(self.tetra, nt, self.pd)
        bm = R.create_bitmap(nt * 4)
        n_tri = R.find_forward_hull(self.tetra, nt, self.pd, n, bm)
        self.tri = R.copy_fh_triangles(self.tetra, nt, self.pd, n, n_tri, bm)
        R.free(bm)

    def map_bands(self, grid_dim, box_len, box_len_z, center, n_samp, delta_sample, seed,
                  band_begin, band_step, rho=None):
        """The reference's pt_loc_2d and shoot_3d_ray over every band_step-th
        4-row band from band_begin (single-threaded); returns (rho, rays)."""
        if rho is None:
            rho = np.zeros(grid_dim * grid_dim)
        rays = self.R.refh_map_bands(self.pd, self.tetra, self.nabla, self.tri, grid_dim,
                                     box_len, box_len_z, rho, center[0], center[1], center[2],
                                     n_samp, delta_sample, seed, band_begin, band_step)
        return rho, int(rays)

    def volume(self, grid_dim, box_len, center, supersampling, seed, grid=None):
        """interpolate_to_3d_grid's grid by the reference's own
        evaluate_3d_ray_on_grid (single-threaded; perturbation corners keyed
        as in the oracle); returns (grid, rays)."""
        if grid is None:
            grid = np.zeros(grid_dim ** 3)
        self.R.refh_set_ray(seed, 0, 0)  # names the stream's seed
        rays = self.R.refh_volume(self.pd, self.tetra, self.nabla, self.tri, grid_dim, box_len,
                                  grid, center[0], center[1], center[2], supersampling)
        return grid, int(rays)

    def close(self):
        if self.nabla:
            self.R.free(C.c_void_p(self.nabla))
            self.R.free(C.c_void_p(self.tri))
            self.nabla = self.tri = None
