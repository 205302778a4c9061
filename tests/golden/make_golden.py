"""Write the golden vectors under tests/golden/ by running the reference C.

Run in a container that holds the reference tree (the reference build is
made from /root/reference by ``make -C oracle ref``):

    python tests/golden/make_golden.py

Every output below comes from the reference's own functions
(pre_compute_vol, gradients, find_forward_hull, copy_fh_triangles, pt_loc_2d,
shoot_3d_ray, compute_density, compute_3d_density, interpolate_to_3d_grid,
evaluate_3d_ray_on_grid) with libc rand() answered from the shared
sample stream (oracle/ref_harness.c).  Inputs are stored beside the outputs;
the one large tessellation (the 32,768-particle map) is stored as a SHA-256
of its bytes and rebuilt by the tests.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from sdtfe_amd import workloads  # noqa: E402
from sdtfe_amd.tessellate import orient_and_link, to_particle_data, triangulate  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


class RefField:
    """Reference density, gradients and forward-hull triangles of one
    tessellation, kept as C pointers for the walk calls."""

    def __init__(self, R, pos, tetra, p_mass=1.0):
        self.R = R
        self.pd = to_particle_data(pos)
        self.tetra = np.ascontiguousarray(tetra, dtype=np.int32)
        n, nt = self.pd.shape[0], self.tetra.shape[0]
        ptr = C.c_void_p(self.pd.ctypes.data)
        R.pre_compute_vol(C.byref(ptr), n, self.tetra, nt, p_mass)
        self.nabla_ptr = R.gradients(self.tetra, nt, self.pd)
        self.nabla = np.ctypeslib.as_array(C.cast(self.nabla_ptr, C.POINTER(C.c_double)),
                                           shape=(nt * 3,)).copy().reshape(nt, 3)
        bm = R.create_bitmap(nt * 4)
        self.n_tri = R.find_forward_hull(self.tetra, nt, self.pd, n, bm)
        self.tri_ptr = R.copy_fh_triangles(self.tetra, nt, self.pd, n, self.n_tri, bm)
        R.free(bm)
        tri = np.ctypeslib.as_array(C.cast(self.tri_ptr, C.POINTER(C.c_int32)),
                                    shape=(self.n_tri * 8,)).copy().reshape(-1, 8)
        tri[:, 7] = 0  # slot 7 is never written by the reference
        self.tri = tri

    def locate(self, x, y):
        start = C.c_int(0)
        return self.R.pt_loc_2d(self.tri_ptr, self.pd, C.byref(start), x, y)

    def shoot(self, entry, x, y, off_z, lz, dxy, seed, pixel, sample):
        self.R.refh_set_ray(seed, pixel, sample)
        crossed = C.c_int(0)
        v = self.R.shoot_3d_ray(entry, self.tetra, self.nabla_ptr, self.pd, x, y, off_z, lz,
                                dxy, 0, 0, C.byref(crossed))
        return v, crossed.value


def single_tet():
    pos = np.array([[0.1, 0.2, 0.05], [0.9, 0.25, 0.1], [0.3, 0.85, 0.2], [0.45, 0.4, 0.9]])
    simp = np.array([[0, 1, 2, 3]])
    nb = -np.ones((1, 4), dtype=np.int64)
    return pos, orient_and_link(pos, simp, nb)


def g1_tiny(R):
    rng = np.random.default_rng(11)
    cases = {"tet4": single_tet()}
    for name, n in (("pts8", 8), ("pts64", 64)):
        p = rng.random((n, 3))
        cases[name] = (p, triangulate(p))
    p = rng.integers(0, 64, size=(512, 3)).astype(np.float64) / 64.0
    p = np.unique(p, axis=0)
    cases["dyadic"] = (p, triangulate(p))
    out = {}
    for name, (p, t) in cases.items():
        f = RefField(R, p, t, p_mass=0.75)
        out[f"{name}_pos"] = p
        out[f"{name}_tetra"] = f.tetra
        out[f"{name}_rho"] = f.pd[:, 3].copy()
        out[f"{name}_nabla"] = f.nabla
        out[f"{name}_tri"] = f.tri
    np.savez_compressed(os.path.join(OUT, "g1_tiny.npz"), **out)


def stream_xy(seed, G, S, dxy, off_x, off_y, scale, soff):
    pix = np.repeat(np.arange(G * G, dtype=np.int64), S)
    smp = np.tile(np.arange(S, dtype=np.int64), G * G)
    r0 = np.array([O.draw31(seed, int(p), int(s), 0) for p, s in zip(pix, smp)], dtype=np.float64)
    r1 = np.array([O.draw31(seed, int(p), int(s), 1) for p, s in zip(pix, smp)], dtype=np.float64)
    px = (pix % G).astype(np.float64)
    py = (pix // G).astype(np.float64)
    x = px * dxy + off_x - soff + (r0 / 2147483647) * scale
    y = py * dxy + off_y - soff + (r1 / 2147483647) * scale
    return pix, smp, x, y


def g2_rays(R):
    rng = np.random.default_rng(1)
    pos = rng.random((2048, 3))
    tetra = triangulate(pos)
    f = RefField(R, pos, tetra)
    seed, G, S = 5, 64, 4
    dxy = 1.0 / G
    pix, smp, x, y = stream_xy(seed, G, S, dxy, 0.0, 0.0, dxy, 0.0)
    # crafted rays: exactly through vertices and along projected edges, which
    # makes Plücker products vanish and drives the perturbation path
    interior = np.where((pos[:, 0] > 0.1) & (pos[:, 0] < 0.9) & (pos[:, 1] > 0.1)
                        & (pos[:, 1] < 0.9))[0][:300]
    vx, vy = pos[interior, 0], pos[interior, 1]
    e = tetra[rng.choice(tetra.shape[0], 300, replace=False)]
    ex = (pos[e[:, 0], 0] + pos[e[:, 1], 0]) / 2
    ey = (pos[e[:, 0], 1] + pos[e[:, 1], 1]) / 2
    n_c = vx.size + ex.size
    cx = np.concatenate([vx, ex])
    cy = np.concatenate([vy, ey])
    cpix = np.arange(n_c, dtype=np.int64) + 10_000_000
    csmp = np.zeros(n_c, dtype=np.int64)
    X = np.concatenate([x, cx])
    Y = np.concatenate([y, cy])
    P = np.concatenate([pix, cpix])
    Sm = np.concatenate([smp, csmp])
    entry = np.empty(X.size, dtype=np.int32)
    val = np.empty(X.size)
    crossed = np.empty(X.size, dtype=np.int32)
    for i in range(X.size):
        entry[i] = f.locate(X[i], Y[i])
        val[i], crossed[i] = f.shoot(int(entry[i]), X[i], Y[i], 0.0, 1.0, dxy, seed,
                                     int(P[i]), int(Sm[i]))
    np.savez_compressed(os.path.join(OUT, "g2_rays.npz"), pos=pos, tetra=f.tetra,
                        rho=f.pd[:, 3].copy(), nabla=f.nabla, tri=f.tri, seed=seed,
                        dxy=dxy, off_z=0.0, lz=1.0, x=X, y=Y, pixel=P, sample=Sm,
                        entry=entry, value=val, crossed=crossed, n_crafted=n_c)
    return int((crossed[-n_c:] > 0).sum())


MAPS = [
    # name, grid, samples, box_len, box_len_z, center, delta_sample, p_mass, seed
    ("plain", 64, 4, 1.0, 1.0, (0.5, 0.5, 0.5), 0.0, 1.0, 3),
    ("slab", 32, 6, 0.8, 0.3, (0.45, 0.55, 0.5), 0.05, 0.37, 11),
    ("wide", 48, 3, 1.4, 1.2, (0.5, 0.5, 0.5), 0.0, 2.5, 17),
]


def g3_maps(R):
    rng = np.random.default_rng(1)
    pos = rng.random((2048, 3))
    tetra = triangulate(pos)
    out = {"pos": pos, "tetra": tetra}
    for name, G, S, L, Lz, cen, ds, m, seed in MAPS:
        pd = to_particle_data(pos)
        rho = np.zeros(G * G)
        R.refh_set_map(seed, G, S)
        R.compute_density(pd, pd.shape[0], tetra, tetra.shape[0], G, L, Lz, rho, m,
                          cen[0], cen[1], cen[2], S, ds)
        assert R.refh_sample_draws() == 2 * G * G * S
        out[f"{name}_rho"] = rho
        out[f"{name}_vertex_rho"] = pd[:, 3].copy()
    np.savez_compressed(os.path.join(OUT, "g3_maps.npz"), **out)


def g4_c1(R):
    pos = workloads.uniform(32768, 0)
    tetra = triangulate(pos)
    spec = workloads.C1["spec"]
    pd = to_particle_data(pos)
    G, S = spec.grid_dim, spec.n_samples
    rho = np.zeros(G * G)
    R.refh_set_map(spec.seed, G, S)
    R.compute_density(pd, pd.shape[0], tetra, tetra.shape[0], G, spec.box_len, spec.box_len_z,
                      rho, spec.p_mass, *spec.center, S, spec.delta_sample)
    np.savez_compressed(os.path.join(OUT, "g4_c1.npz"), rho=rho, pos_sha=sha(pos),
                        tetra_sha=sha(tetra), n_tetra=tetra.shape[0],
                        vertex_rho_sum=float(pd[:, 3].sum()), seed=spec.seed)


def g5_utils(R):
    """convert_to_qhdata and rotate3d (src/utils.c:8-16, 33-79), the host
    pre-processing of the pydtfe binding."""
    rng = np.random.default_rng(21)
    pos32 = rng.random((500, 3)).astype(np.float32)
    planar = np.ascontiguousarray(pos32.T).ravel()  # Fortran order of (N, 3)
    ptr = R.convert_to_qhdata(planar, 500)
    qh = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_double)), shape=(2000,)).copy()
    R.free(ptr)
    qh = qh.reshape(500, 4)
    theta = np.array([0.3, -1.1, 2.5])
    center = np.array([0.5, 0.4, 0.6])
    rot = qh.copy()
    R.rotate3d(rot, 500, theta, center)
    np.savez_compressed(os.path.join(OUT, "g5_utils.npz"), pos32=pos32, qh=qh[:, :3],
                        theta=theta, center=center, rotated=rot[:, :3])


VOLUMES = [
    # name, grid, supersampling, box_len, center, p_mass (None: given field), seed
    ("dens", 16, 2, 1.0, (0.5, 0.5, 0.5), 0.37, 0),
    ("field", 12, 3, 1.3, (0.45, 0.55, 0.5), None, 0),
]


def _field(pos):
    """A vertex field that is not linear (so gradients vary per tet)."""
    return 1.0 + 2.0 * pos[:, 0] - 3.0 * pos[:, 1] + 0.5 * pos[:, 2] + np.sin(7.0 * pos[:, 2])


def _volume_by_rays(R, pd, tetra, G, ss, L, cen, seed, grid):
    """interpolate_to_3d_grid's column loop (src/dtfe.c:686-731) driven from
    here so that each ray's perturbation draws are keyed (seed, column, ray):
    the reference's own pt_loc_2d and evaluate_3d_ray_on_grid per ray.
    Returns the entry tets."""
    nt = tetra.shape[0]
    nabla_ptr = R.gradients(tetra, nt, pd)
    bm = R.create_bitmap(nt * 4)
    n_tri = R.find_forward_hull(tetra, nt, pd, pd.shape[0], bm)
    tri_ptr = R.copy_fh_triangles(tetra, nt, pd, pd.shape[0], n_tri, bm)
    R.free(bm)
    x, y = O.volume_ray_xy(G, L, cen, ss)
    off_z = -(L / 2.0) + cen[2]
    dxy = L / G
    entries = np.empty(x.size, dtype=np.int32)
    tmp = C.c_int(0)
    for q in range(x.size):
        c, r = divmod(q, ss * ss)
        start = C.c_int(0)
        entries[q] = R.pt_loc_2d(tri_ptr, pd, C.byref(start), float(x[q]), float(y[q]))
        R.refh_set_ray(seed, c, r)
        R.evaluate_3d_ray_on_grid(int(entries[q]), tetra, nabla_ptr, pd, float(x[q]),
                                  float(y[q]), off_z, L, G, dxy, C.byref(tmp), ss,
                                  grid.ctypes.data + 8 * c * G)
    R.free(C.c_void_p(nabla_ptr))
    R.free(C.c_void_p(tri_ptr))
    return entries


def g6_volume(R):
    """Volumetric grids (src/dtfe.c:376-516, 657-735) from the reference."""
    rng = np.random.default_rng(1)
    pos = rng.random((2048, 3))
    tetra = triangulate(pos)
    out = {"pos": pos, "tetra": tetra}
    for name, G, ss, L, cen, m, seed in VOLUMES:
        pd = to_particle_data(pos)
        init = np.zeros(G ** 3) if m is not None else rng.random(G ** 3)
        grid = init.copy()
        R.refh_reset_rand_calls()
        if m is not None:
            R.compute_3d_density(pd, pd.shape[0], tetra, tetra.shape[0], G, L, grid, m,
                                 cen[0], cen[1], cen[2], ss)
            out[f"{name}_vertex_rho"] = pd[:, 3].copy()
        else:
            pd[:, 3] = _field(pos)
            R.interpolate_to_3d_grid(pd, pd.shape[0], tetra, tetra.shape[0], G, L, grid,
                                     cen[0], cen[1], cen[2], ss)
            # the same grid from the per-ray driver (pins its loop restatement)
            by_rays = init.copy()
            _volume_by_rays(R, pd, tetra, G, ss, L, cen, seed, by_rays)
            assert np.array_equal(by_rays, grid)
        assert R.refh_rand_calls() == 0  # no perturbation: draws play no part
        out[f"{name}_init"] = init
        out[f"{name}_grid"] = grid
    # rays exactly through vertices and along edges of a dyadic point set,
    # which drives the perturbation (keyed draws through the per-ray driver)
    p = np.unique(rng.integers(0, 64, size=(700, 3)).astype(np.float64) / 64.0, axis=0)
    t = triangulate(p)
    pd = to_particle_data(p)
    R.pre_compute_vol(C.byref(C.c_void_p(pd.ctypes.data)), pd.shape[0], t, t.shape[0], 1.0)
    for name, G, ss in (("crafted16", 16, 1), ("crafted8", 8, 2)):
        grid = np.zeros(G ** 3)
        R.refh_reset_rand_calls()
        entries = _volume_by_rays(R, pd, t, G, ss, 1.0, (0.5, 0.5, 0.5), 7, grid)
        out[f"{name}_grid"] = grid
        out[f"{name}_entries"] = entries
        out[f"{name}_rand_calls"] = R.refh_rand_calls()
    out["crafted_pos"] = p
    out["crafted_tetra"] = t
    out["crafted_rho"] = pd[:, 3].copy()
    np.savez_compressed(os.path.join(OUT, "g6_volume.npz"), **out)
    return {k: int(v) for k, v in out.items() if k.endswith("_rand_calls")}


def main():
    if not O.ref_available():
        O.build(ref=True)
    O.build()
    R = O.ref()
    g1_tiny(R)
    hit = g2_rays(R)
    g3_maps(R)
    g4_c1(R)
    g5_utils(R)
    pert = g6_volume(R)
    print("golden vectors written; crafted rays that completed:", hit,
          "; volumetric perturbation draws:", pert)


if __name__ == "__main__":
    main()
