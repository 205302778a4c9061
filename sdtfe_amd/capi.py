"""ctypes binding of libsdtfe_amd.so (the C ABI in include/sdtfe_amd.h).

The library is the product path; there is no CPU fallback.  Loading fails
loudly if the library has not been built (``python -m sdtfe_amd.build``).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsdtfe_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "sdtfe_amd.h")

PREC_FP64 = 0
PREC_MIXED = 1
MAX_STEPS_DEFAULT = 1 << 24


class SdtfeError(RuntimeError):
    pass


class MapParams(C.Structure):
    _fields_ = [("grid_dim", C.c_int), ("n_mc_samp", C.c_int), ("box_len", C.c_double),
                ("box_len_z", C.c_double), ("center", C.c_double * 3),
                ("delta_sample", C.c_double), ("seed", C.c_uint64), ("precision", C.c_int),
                ("max_steps", C.c_int)]


class Stats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "samples", "missed", "failed", "no_intersection", "perturb_exhausted",
        "no_isect_perturb", "perturb_steps", "tet_steps", "step_cap", "fp64_redo")] + [
        ("t_walk_ms", C.c_double), ("t_total_ms", C.c_double), ("t_kernel_ms", C.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class TessInfo(C.Structure):
    _fields_ = [("n_particles", C.c_int), ("n_tetra", C.c_int), ("n_tri", C.c_int),
                ("grid_nx", C.c_int), ("grid_ny", C.c_int), ("device_bytes", C.c_int64),
                ("t_setup_ms", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_int64), ("max_active", C.c_int),
                ("tile_budget", C.c_int), ("lanes_per_pixel", C.c_int), ("precision", C.c_int),
                ("setup_threads", C.c_int)]


class RequestInfo(C.Structure):
    _fields_ = [("status", C.c_int), ("queue_ms", C.c_double), ("setup_ms", C.c_double),
                ("latency_ms", C.c_double), ("submit_ms", C.c_double), ("done_ms", C.c_double),
                ("footprint_bytes", C.c_int64), ("stats", Stats)]


class ServerInfo(C.Structure):
    _fields_ = [("queued", C.c_int), ("active", C.c_int), ("used_bytes", C.c_int64),
                ("launches", C.c_int64), ("walk_ms", C.c_double)]


class VolumeParams(C.Structure):
    _fields_ = [("grid_dim", C.c_int), ("supersampling", C.c_uint), ("box_len", C.c_double),
                ("center", C.c_double * 3), ("seed", C.c_uint64), ("max_steps", C.c_int)]


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_vp = C.c_void_p

_SIGS = {
    "sdtfe_last_error": (C.c_char_p, []),
    "sdtfe_set_seed": (None, [C.c_uint64]),
    "sdtfe_get_seed": (C.c_uint64, []),
    "sdtfe_tess_create": (C.c_int, [_dp, C.c_int, _ip, C.c_int, C.c_float, C.POINTER(_vp)]),
    "sdtfe_tess_create_field": (C.c_int, [_dp, C.c_int, _ip, C.c_int, C.POINTER(_vp)]),
    "sdtfe_tess_destroy": (None, [_vp]),
    "sdtfe_tess_get_info": (C.c_int, [_vp, C.POINTER(TessInfo)]),
    "sdtfe_tess_vertex_density": (C.c_int, [_vp, _dp]),
    "sdtfe_tess_gradients": (C.c_int, [_vp, _dp]),
    "sdtfe_tess_triangles": (C.c_int, [_vp, _ip, C.c_int]),
    "sdtfe_tess_walk_records": (C.c_int, [_vp, _vp, C.c_size_t]),
    "sdtfe_tess_locate": (C.c_int, [_vp, C.c_int, _dp, _dp, _ip]),
    "sdtfe_tess_trace": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp, _vp, _u64p, _u32p,
                                   C.c_double, C.c_double, C.c_double, C.c_uint64, _dp, _ip,
                                   C.POINTER(Stats)]),
    "sdtfe_map": (C.c_int, [_vp, C.POINTER(MapParams), _dp, _vp, C.POINTER(Stats)]),
    "sdtfe_map_device": (C.c_int, [_vp, C.POINTER(MapParams), _vp, C.POINTER(Stats)]),
    "sdtfe_map_device_noalloc": (C.c_int, [_vp, C.POINTER(MapParams), _vp, C.POINTER(Stats)]),
    "sdtfe_tess_sync": (C.c_int, [_vp]),
    "sdtfe_tetra_from_simplices": (C.c_int, [_dp, C.c_int, _ip, C.c_int, _ip]),
    "sdtfe_volume": (C.c_int, [_vp, C.POINTER(VolumeParams), _dp, C.POINTER(Stats)]),
    "sdtfe_volume_device": (C.c_int, [_vp, C.POINTER(VolumeParams), _vp, C.POINTER(Stats)]),
    "sdtfe_batch_create": (C.c_int, [C.POINTER(_vp)]),
    "sdtfe_batch_destroy": (None, [_vp]),
    "sdtfe_batch_run": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), C.POINTER(MapParams),
                                  C.POINTER(_vp), C.POINTER(C.c_int), C.POINTER(Stats),
                                  C.POINTER(Stats)]),
    "sdtfe_server_create": (C.c_int, [C.POINTER(ServerConfig), C.POINTER(_vp)]),
    "sdtfe_server_submit": (C.c_int, [_vp, _dp, C.c_int, _ip, C.c_int, C.c_float,
                                      C.POINTER(MapParams), _dp, C.POINTER(C.c_int64)]),
    "sdtfe_server_wait": (C.c_int, [_vp, C.c_int64, C.POINTER(RequestInfo)]),
    "sdtfe_server_get_info": (C.c_int, [_vp, C.POINTER(ServerInfo)]),
    "sdtfe_server_destroy": (None, [_vp]),
    "sdtfe_device_alloc": (_vp, [C.c_size_t]),
    "sdtfe_device_free": (C.c_int, [_vp]),
    "sdtfe_memcpy_d2h": (C.c_int, [_vp, _vp, C.c_size_t]),
    "sdtfe_memcpy_h2d": (C.c_int, [_vp, _vp, C.c_size_t]),
    "sdtfe_memset_d": (C.c_int, [_vp, C.c_int, C.c_size_t]),
    "sdtfe_device_count": (C.c_int, []),
    "compute_density": (None, [_dp, C.c_int, _ip, C.c_int, C.c_int, C.c_double, C.c_double,
                               _dp, C.c_float, C.c_double, C.c_double, C.c_double, C.c_int,
                               C.c_double]),
    "pre_compute_vol": (None, [C.POINTER(_vp), C.c_int, _ip, C.c_int, C.c_float]),
    "gradients": (_vp, [_ip, C.c_int, _dp]),
    "interpolate_to_3d_grid": (None, [_dp, C.c_int, _ip, C.c_int, C.c_int, C.c_double, _dp,
                                      C.c_double, C.c_double, C.c_double, C.c_uint]),
    "compute_3d_density": (None, [_dp, C.c_int, _ip, C.c_int, C.c_int, C.c_double, _dp,
                                  C.c_float, C.c_double, C.c_double, C.c_double, C.c_uint]),
    "sdtfe_dropin_status": (C.c_int, []),
    "find_forward_hull": (C.c_int, [_ip, C.c_int, _dp, C.c_int, _u32p]),
    "copy_fh_triangles": (_vp, [_ip, C.c_int, _dp, C.c_int, C.c_int, _u32p]),
    "convert_to_qhdata": (_vp, [_f32p, C.c_size_t]),
    "rotate3d": (None, [_dp, C.c_size_t, _dp, _dp]),
    "shuffle_buff": (None, [_dp, C.c_int, C.c_int]),
    "set_bit": (None, [_u32p, C.c_int]),
    "clear_bit": (None, [_u32p, C.c_int]),
    "get_bit": (C.c_int, [_u32p, C.c_int]),
    "create_bitmap": (_vp, [C.c_int]),
    "delete_bitmap": (None, [_vp]),
    "write_image": (None, [C.c_char_p, _dp, C.c_int]),
    "write_rho": (C.c_int, [C.c_char_p, C.POINTER(_vp), C.c_int]),
}

_lib = None


def lib():
    """Load libsdtfe_amd.so; raise if it is missing (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SdtfeError(f"{LIB_PATH} not built; run `python -m sdtfe_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def header_functions() -> list[str]:
    """Function names declared in include/sdtfe_amd.h."""
    import re
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", txt)))


def check(rc: int) -> None:
    if rc != 0:
        raise SdtfeError(f"sdtfe error {rc}: {lib().sdtfe_last_error().decode()}")


def check_dropin(name: str) -> None:
    """Raise SdtfeError if the calling thread's last drop-in call failed (the
    reference's void prototypes carry no status; the library keeps it)."""
    rc = int(lib().sdtfe_dropin_status())
    if rc != 0:
        raise SdtfeError(f"{name} failed ({rc}): {lib().sdtfe_last_error().decode()}")


def device_count() -> int:
    return int(lib().sdtfe_device_count())


def _c64(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


class DeviceBuffer:
    """A device allocation owned by Python (freed on close/GC)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = lib().sdtfe_device_alloc(self.nbytes)
        if not self.ptr:
            raise SdtfeError(lib().sdtfe_last_error().decode())

    def zero(self):
        check(lib().sdtfe_memset_d(self.ptr, 0, self.nbytes))

    def to_host(self, dtype=np.float64, count=None):
        n = self.nbytes // np.dtype(dtype).itemsize if count is None else count
        out = np.empty(n, dtype=dtype)
        check(lib().sdtfe_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def close(self):
        if self.ptr:
            lib().sdtfe_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Tessellation:
    """A tessellation resident in HBM (sdtfe_tess)."""

    def __init__(self, particle_data: np.ndarray, tetra_data: np.ndarray, p_mass: float = 1.0,
                 given_field: bool = False):
        pd = _c64(particle_data)
        if pd.ndim != 2 or pd.shape[1] != 4:
            raise ValueError("particle_data must be N x 4 float64")
        tet = _c64(tetra_data, np.int32)
        if tet.ndim != 2 or tet.shape[1] != 8:
            raise ValueError("tetra_data must be T x 8 int32")
        h = C.c_void_p()
        if given_field:
            rc = lib().sdtfe_tess_create_field(pd, pd.shape[0], tet, tet.shape[0], C.byref(h))
        else:
            rc = lib().sdtfe_tess_create(pd, pd.shape[0], tet, tet.shape[0], p_mass, C.byref(h))
        check(rc)
        self.h = h
        self.n = pd.shape[0]
        self.nt = tet.shape[0]

    def close(self):
        if getattr(self, "h", None):
            lib().sdtfe_tess_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def info(self) -> dict:
        i = TessInfo()
        check(lib().sdtfe_tess_get_info(self.h, C.byref(i)))
        return {n: getattr(i, n) for n, _ in i._fields_}

    def vertex_density(self) -> np.ndarray:
        out = np.empty(self.n)
        check(lib().sdtfe_tess_vertex_density(self.h, out))
        return out

    def gradients(self) -> np.ndarray:
        out = np.empty((self.nt, 3))
        check(lib().sdtfe_tess_gradients(self.h, out))
        return out

    def triangles(self) -> np.ndarray:
        f = self.info()["n_tri"]
        out = np.zeros((max(f, 1), 8), dtype=np.int32)
        check(lib().sdtfe_tess_triangles(self.h, out, f))
        return out[:f]

    REC64 = np.dtype([("nb", "<i4", 4), ("v", "<f8", 12), ("rho_a", "<f8"), ("g", "<f8", 3)])

    def walk_records(self) -> np.ndarray:
        out = np.zeros(self.nt, dtype=self.REC64)
        check(lib().sdtfe_tess_walk_records(self.h, out.ctypes.data, out.nbytes))
        return out

    def locate(self, x, y) -> np.ndarray:
        x, y = _c64(x), _c64(y)
        out = np.empty(x.size, dtype=np.int32)
        check(lib().sdtfe_tess_locate(self.h, x.size, x, y, out))
        return out

    def trace(self, x, y, pixel, sample, off_z, lz, delta_xy, seed, entry=None,
              precision=PREC_FP64):
        x, y = _c64(x), _c64(y)
        n = x.size
        pix = _c64(pixel, np.uint64)
        smp = _c64(sample, np.uint32)
        val = np.empty(n)
        crossed = np.empty(n, dtype=np.int32)
        ent = None if entry is None else _c64(entry, np.int32)
        st = Stats()
        check(lib().sdtfe_tess_trace(self.h, precision, n, x, y,
                                     None if ent is None else ent.ctypes.data, pix, smp,
                                     off_z, lz, delta_xy, seed, val, crossed, C.byref(st)))
        return val, crossed, st.as_dict()

    def params(self, grid_dim, n_samples, box_len=1.0, box_len_z=1.0, center=(0.5, 0.5, 0.5),
               delta_sample=0.0, seed=0, precision=PREC_FP64, max_steps=0) -> MapParams:
        p = MapParams()
        p.grid_dim, p.n_mc_samp = int(grid_dim), int(n_samples)
        p.box_len, p.box_len_z = float(box_len), float(box_len_z)
        p.center[:] = [float(c) for c in center]
        p.delta_sample, p.seed = float(delta_sample), int(seed)
        p.precision, p.max_steps = int(precision), int(max_steps)
        return p

    def map(self, grid_dim, n_samples, rho=None, want_navg=False, **kw):
        """One map on the GPU; returns (rho, stats[, n_avg]).  rho is
        accumulated into if given."""
        p = self.params(grid_dim, n_samples, **kw)
        npix = grid_dim * grid_dim
        out = np.zeros(npix) if rho is None else rho
        if out.dtype != np.float64 or not out.flags.c_contiguous or out.size != npix:
            raise ValueError("rho must be a contiguous float64 array of grid_dim^2")
        navg = np.empty(npix, dtype=np.int32) if want_navg else None
        st = Stats()
        check(lib().sdtfe_map(self.h, C.byref(p), out,
                              None if navg is None else navg.ctypes.data, C.byref(st)))
        return (out, st.as_dict(), navg) if want_navg else (out, st.as_dict())

    def map_device(self, params: MapParams, rho_dev: DeviceBuffer, stats: bool = True):
        st = Stats() if stats else None
        check(lib().sdtfe_map_device(self.h, C.byref(params), rho_dev.ptr,
                                     C.byref(st) if st is not None else None))
        return st.as_dict() if st is not None else None

    def map_device_noalloc(self, params: MapParams, rho_dev: DeviceBuffer, stats: bool = True):
        """sdtfe_map_device without allocation (the library's scratch memory)."""
        st = Stats() if stats else None
        check(lib().sdtfe_map_device_noalloc(self.h, C.byref(params), rho_dev.ptr,
                                             C.byref(st) if st is not None else None))
        return st.as_dict() if st is not None else None

    @staticmethod
    def volume_params(grid_dim, supersampling=1, box_len=1.0, center=(0.5, 0.5, 0.5), seed=0,
                      max_steps=0) -> VolumeParams:
        p = VolumeParams()
        p.grid_dim, p.supersampling = int(grid_dim), int(supersampling)
        p.box_len = float(box_len)
        p.center[:] = [float(c) for c in center]
        p.seed, p.max_steps = int(seed), int(max_steps)
        return p

    def volume(self, grid_dim, supersampling=1, grid=None, **kw):
        """The volumetric grid of the vertex field on the GPU
        (interpolate_to_3d_grid); returns (grid, stats), grid flat
        grid_dim^3 in [x][y][z] order, accumulated into if given."""
        p = self.volume_params(grid_dim, supersampling, **kw)
        cells = grid_dim ** 3
        out = np.zeros(cells) if grid is None else grid
        if out.dtype != np.float64 or not out.flags.c_contiguous or out.size != cells:
            raise ValueError("grid must be a contiguous float64 array of grid_dim^3")
        st = Stats()
        check(lib().sdtfe_volume(self.h, C.byref(p), out, C.byref(st)))
        return out, st.as_dict()

    def volume_device(self, params: VolumeParams, grid_dev: "DeviceBuffer", stats: bool = True):
        st = Stats() if stats else None
        check(lib().sdtfe_volume_device(self.h, C.byref(params), grid_dev.ptr,
                                        C.byref(st) if st is not None else None))
        return st.as_dict() if st is not None else None

    def sync(self):
        check(lib().sdtfe_tess_sync(self.h))


class Batch:
    """Many maps in one persistent, graph-replayed launch (sdtfe_batch)."""

    def __init__(self):
        h = C.c_void_p()
        check(lib().sdtfe_batch_create(C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().sdtfe_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def run(self, items, lanes_per_pixel=None):
        """items: list of (Tessellation, MapParams, DeviceBuffer).  Returns
        (per-item stats, total stats); maps accumulate into the buffers."""
        n = len(items)
        tess = (C.c_void_p * n)(*[it[0].h.value for it in items])
        params = (MapParams * n)(*[it[1] for it in items])
        outs = (C.c_void_p * n)(*[it[2].ptr for it in items])
        lpp = None
        if lanes_per_pixel is not None:
            lpp = (C.c_int * n)(*[int(v) for v in lanes_per_pixel])
        st = (Stats * n)()
        tot = Stats()
        check(lib().sdtfe_batch_run(self.h, n, tess, params, outs, lpp, st, C.byref(tot)))
        return [x.as_dict() for x in st], tot.as_dict()


class Server:
    """Map server with continuous batching (sdtfe_server)."""

    def __init__(self, hbm_budget_bytes=64 << 30, max_active=64, tile_budget=4096,
                 lanes_per_pixel=0, precision=PREC_FP64, setup_threads=0):
        cfg = ServerConfig(int(hbm_budget_bytes), int(max_active), int(tile_budget),
                           int(lanes_per_pixel), int(precision), int(setup_threads))
        h = C.c_void_p()
        check(lib().sdtfe_server_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.precision = precision
        self._out = {}  # request id -> output array kept alive until waited

    def submit(self, particle_data, tetra_data, params: MapParams, p_mass=1.0, rho=None):
        pd = _c64(particle_data)
        tet = _c64(tetra_data, np.int32)
        npix = params.grid_dim * params.grid_dim
        out = np.zeros(npix) if rho is None else rho
        rid = C.c_int64()
        check(lib().sdtfe_server_submit(self.h, pd, pd.shape[0], tet, tet.shape[0], p_mass,
                                        C.byref(params), out, C.byref(rid)))
        self._out[rid.value] = out
        return rid.value

    def wait(self, rid):
        info = RequestInfo()
        rc = lib().sdtfe_server_wait(self.h, rid, C.byref(info))
        out = self._out.pop(rid)
        check(rc)
        d = {n: getattr(info, n) for n, _ in info._fields_ if n != "stats"}
        d["stats"] = info.stats.as_dict()
        return out, d

    def info(self):
        i = ServerInfo()
        check(lib().sdtfe_server_get_info(self.h, C.byref(i)))
        return {n: getattr(i, n) for n, _ in i._fields_}

    def close(self):
        if getattr(self, "h", None):
            lib().sdtfe_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def compute_density(particle_data, tetra_data, grid_dim, box_len, box_len_z, rho, p_mass,
                    center, n_mc_samp, delta_sample) -> None:
    """The drop-in C entry point (src/dtfe.h:13-15) called through the ABI:
    particle_data[:, 3] is overwritten and rho accumulated in place."""
    if particle_data.dtype != np.float64 or not particle_data.flags.c_contiguous:
        raise ValueError("particle_data must be C-contiguous float64")
    tet = _c64(tetra_data, np.int32)
    lib().compute_density(particle_data, particle_data.shape[0], tet, tet.shape[0], grid_dim,
                          box_len, box_len_z, rho, p_mass, center[0], center[1], center[2],
                          n_mc_samp, delta_sample)
    check_dropin("compute_density")


def interpolate_to_3d_grid(particle_data, tetra_data, grid_dim, box_len, grid, center,
                           supersampling) -> None:
    """The drop-in C entry point (src/dtfe.h:17-18) through the ABI: the
    field in particle_data[:, 3] is sampled into grid (accumulated)."""
    if particle_data.dtype != np.float64 or not particle_data.flags.c_contiguous:
        raise ValueError("particle_data must be C-contiguous float64")
    tet = _c64(tetra_data, np.int32)
    lib().interpolate_to_3d_grid(particle_data, particle_data.shape[0], tet, tet.shape[0],
                                 grid_dim, box_len, grid, center[0], center[1], center[2],
                                 supersampling)
    check_dropin("interpolate_to_3d_grid")


def compute_3d_density(particle_data, tetra_data, grid_dim, box_len, rho, p_mass, center,
                       supersampling) -> None:
    """The drop-in C entry point (src/dtfe.h:20-21) through the ABI:
    particle_data[:, 3] is overwritten and rho accumulated in place."""
    if particle_data.dtype != np.float64 or not particle_data.flags.c_contiguous:
        raise ValueError("particle_data must be C-contiguous float64")
    tet = _c64(tetra_data, np.int32)
    lib().compute_3d_density(particle_data, particle_data.shape[0], tet, tet.shape[0],
                             grid_dim, box_len, rho, p_mass, center[0], center[1], center[2],
                             supersampling)
    check_dropin("compute_3d_density")


def tetra_from_simplices(particle_data, simplices) -> np.ndarray:
    """tetra_data (T x 8 int32, the reference layout) from Delaunay
    simplices (T x 4) and particle_data (N x 4 or N x 3 float64), assembled
    on the GPU: orientation and neighbour slots as triangulate makes them."""
    p = np.asarray(particle_data, dtype=np.float64)
    if p.ndim != 2 or p.shape[1] not in (3, 4):
        raise ValueError("particle_data must be N x 3 or N x 4")
    if p.shape[1] == 3:
        p = np.concatenate([p, np.zeros((p.shape[0], 1))], axis=1)
    p = np.ascontiguousarray(p)
    simp = _c64(simplices, np.int32)
    if simp.ndim != 2 or simp.shape[1] != 4:
        raise ValueError("simplices must be T x 4")
    out = np.empty((simp.shape[0], 8), dtype=np.int32)
    check(lib().sdtfe_tetra_from_simplices(p, p.shape[0], simp, simp.shape[0], out))
    return out


def _libc():
    return C.CDLL(None)


def _take_malloced(ptr, count, ctype, dtype):
    """Copy a malloc'd C array into numpy and free it."""
    if not ptr:
        raise SdtfeError(f"allocation failed: {lib().sdtfe_last_error().decode()}")
    try:
        return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), shape=(count,)).astype(
            dtype, copy=True)
    finally:
        _libc().free(C.c_void_p(ptr))


def find_forward_hull(tetra_data, particle_data, bitmap=None):
    """find_forward_hull (src/triangulate.h:28) through the ABI: returns
    (count, bitmap), bitmap in the reference's layout (bit f of word t for
    face f of tet t; n_tetra + 1 uint32 words)."""
    tet = _c64(tetra_data, np.int32)
    pd = _c64(particle_data)
    if bitmap is None:
        bitmap = np.zeros(tet.shape[0] + 1, dtype=np.uint32)
    n = lib().find_forward_hull(tet, tet.shape[0], pd, pd.shape[0], bitmap)
    check_dropin("find_forward_hull")
    return int(n), bitmap


def copy_fh_triangles(tetra_data, particle_data, n_tri, bitmap) -> np.ndarray:
    """copy_fh_triangles (src/triangulate.h:29) through the ABI: n_tri x 8
    int32 rows (a, b, c, n_CB, n_AC, n_BA, tet, 0)."""
    tet = _c64(tetra_data, np.int32)
    pd = _c64(particle_data)
    bm = np.ascontiguousarray(bitmap, dtype=np.uint32)
    ptr = lib().copy_fh_triangles(tet, tet.shape[0], pd, pd.shape[0], int(n_tri), bm)
    rows = _take_malloced(ptr, 8 * max(int(n_tri), 1), C.c_int, np.int32)
    check_dropin("copy_fh_triangles")
    return rows[:8 * int(n_tri)].reshape(int(n_tri), 8)


def convert_to_qhdata(planar_xyz) -> np.ndarray:
    """convert_to_qhdata (src/utils.h:10): planar float32 (x..., y..., z...)
    to N x 4 float64 rows."""
    buf = np.ascontiguousarray(planar_xyz, dtype=np.float32).ravel()
    n = buf.size // 3
    ptr = lib().convert_to_qhdata(buf, n)
    return _take_malloced(ptr, 4 * max(n, 1), C.c_double, np.float64)[:4 * n].reshape(n, 4)


def rotate3d(data: np.ndarray, n: int, theta, center) -> None:
    """rotate3d (src/utils.h:11) in place on rows 0..n-1 of N x 4 float64."""
    if data.dtype != np.float64 or not data.flags.c_contiguous or data.shape[1] != 4:
        raise ValueError("data must be C-contiguous N x 4 float64")
    lib().rotate3d(data, int(n), _c64(theta), _c64(center))


def shuffle_buff(data: np.ndarray, n_in_buff: int, n_shuffle: int) -> None:
    """shuffle_buff (src/utils.h:12) in place, drawing from the C library's
    rand() (seed it with libc srand)."""
    if data.dtype != np.float64 or not data.flags.c_contiguous or data.shape[1] != 4:
        raise ValueError("data must be C-contiguous N x 4 float64")
    lib().shuffle_buff(data, int(n_in_buff), int(n_shuffle))


def write_image(path: str, rho) -> None:
    """write_image (src/io.h:35): the map as an 8-bit palette TIFF."""
    r = _c64(rho).ravel()
    g = int(round(r.size ** 0.5))
    if g * g != r.size:
        raise ValueError("rho must hold grid_dim^2 values")
    lib().write_image(os.fsencode(path), r, g)


def write_rho(path: str, rho) -> int:
    """write_rho (src/io.h:36): the raw float64 map; returns the count written."""
    r = _c64(rho).ravel()
    ptr = C.c_void_p(r.ctypes.data)
    return int(lib().write_rho(os.fsencode(path), C.byref(ptr), r.size))


def set_seed(seed: int) -> None:
    lib().sdtfe_set_seed(int(seed))
