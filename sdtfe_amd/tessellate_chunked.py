"""Delaunay tetrahedralisation of large particle sets in parallel x-slabs.

Qhull (scipy's copy) is single-threaded; for millions of particles it takes
many minutes.  The Delaunay tetrahedralisation of points in general position
is unique, so it can be assembled from slabs:

* the points are cut into slabs of equal count along x; each slab is
  tetrahedralised together with the points within a margin on both sides
  and the points near the bounding box's y and z faces (the hull's facets
  there can be long in x, and the tets under them reach far);
* a slab keeps the tets whose centroid lies in its own x range, so every tet
  has one owner;
* a kept tet is Delaunay for all points if its circumsphere lies within the
  slab's extended x range (every point that could be inside it was in the
  slab); otherwise its circumsphere is checked against all points with a
  KD-tree, and it is dropped if any other point lies inside;
* the union must then tile the convex hull: every face is shared by at most
  two tets, and the tets' volumes add up to the hull's.  A triangulation of
  the hull whose tets all have empty circumspheres is the Delaunay one.

The result is converted to the reference layout by orient_and_link (the
reference's orientation and neighbour-slot conventions); only the order of
the tets differs from one Qhull run over all points.
"""
from __future__ import annotations

import multiprocessing as mp
import os
import time

import numpy as np

from .tessellate import QHULL_OPTIONS, orient_and_link

_POS = None   # the points, shared with forked workers
_TREE = None  # KD-tree over them


def circumspheres(p: np.ndarray, simp: np.ndarray):
    """Centres, squared radii and 6 x signed volumes of tets simp (T x 4)."""
    a = p[simp[:, 0], :3]
    ba = p[simp[:, 1], :3] - a
    ca = p[simp[:, 2], :3] - a
    da = p[simp[:, 3], :3] - a
    cd = np.cross(ca, da)
    db = np.cross(da, ba)
    bc = np.cross(ba, ca)
    det = np.einsum("ij,ij->i", ba, cd)
    num = (np.einsum("ij,ij->i", ba, ba)[:, None] * cd +
           np.einsum("ij,ij->i", ca, ca)[:, None] * db +
           np.einsum("ij,ij->i", da, da)[:, None] * bc)
    with np.errstate(divide="ignore", invalid="ignore"):
        off = num / (2.0 * det[:, None])
    return a + off, np.einsum("ij,ij->i", off, off), det


def _slab(args):
    from scipy.spatial import Delaunay
    k, idx, lo, hi, ext_lo, ext_hi, first, last = args
    p = _POS
    tri = Delaunay(p[idx, :3], qhull_options=QHULL_OPTIONS)
    simp = idx[tri.simplices]
    cx = p[simp, 0].mean(axis=1)
    own = np.ones(simp.shape[0], dtype=bool)
    if not first:
        own &= cx >= lo
    if not last:
        own &= cx < hi
    simp = simp[own]
    c, r2, _ = circumspheres(p, simp)
    r = np.sqrt(r2)
    fin = np.isfinite(r)
    cert = fin & ((c[:, 0] - r >= ext_lo) | first) & ((c[:, 0] + r <= ext_hi) | last)
    keep = cert.copy()
    unc = np.nonzero(~cert)[0]
    for i in unc:
        keep[i] = _empty_sphere(p, simp[i], c[i], r[i]) if fin[i] else False
    return k, simp[keep], int(unc.size), int((~keep).sum())


def _empty_sphere(p, v, c, r) -> bool:
    """No point strictly inside the circumsphere of tet v: candidates from
    the KD-tree (radius padded), decided by the in-sphere determinant with a
    relative tolerance (points on the sphere do not count)."""
    cand = [j for j in _TREE.query_ball_point(c, r * (1.0 + 1e-9)) if j not in v]
    if not cand:
        return True
    a, b, cc, d = (p[j, :3] for j in v)
    orient = np.dot(b - a, np.cross(cc - a, d - a))
    for j in cand:
        q = p[j, :3]
        rows = np.array([np.append(x - q, np.dot(x - q, x - q)) for x in (a, b, cc, d)])
        det = np.linalg.det(rows)
        scale = np.prod(np.abs(rows).max(axis=1)) + 1e-300
        if det * np.sign(orient) < -1e-10 * scale:  # q inside (sign convention below)
            return False
    return True


def _neighbors(simp: np.ndarray) -> np.ndarray:
    """neighbors[t, k]: the tet across the face opposite vertex k (-1 on the
    hull); raises if a face has more than two tets."""
    t = simp.shape[0]
    s = np.sort(simp, axis=1).astype(np.int64)  # the packed keys need 64 bits
    bits = max(1, int(s.max()).bit_length())
    # face opposite vertex k of the sorted row, as two keys (the first two
    # vertices packed, the third); rows keep their original slot order
    hi = np.empty(4 * t, dtype=np.int64)
    lo = np.empty(4 * t, dtype=np.int64)
    slot = np.empty(4 * t, dtype=np.int64)
    for k in range(4):
        f = np.delete(s, k, axis=1)
        hi[k::4] = (f[:, 0] << bits) | f[:, 1]
        lo[k::4] = f[:, 2]
    # which original vertex slot the sorted column k came from
    src = np.argsort(simp, axis=1)
    for k in range(4):
        slot[k::4] = src[:, k]
    order = np.lexsort((lo, hi))
    h, l = hi[order], lo[order]
    same = (h[1:] == h[:-1]) & (l[1:] == l[:-1])
    if np.any(same[1:] & same[:-1]):
        raise ValueError("a face is shared by more than two tets")
    nb = np.full((t, 4), -1, dtype=np.int64)
    i = np.nonzero(same)[0]
    a, b = order[i], order[i + 1]
    nb[a // 4, slot[a]] = b // 4
    nb[b // 4, slot[b]] = a // 4
    return nb


def triangulate_chunked(pos: np.ndarray, slabs: int = 16, workers: int = 16,
                        margin: float | None = None, shell: float | None = None,
                        log=print, gpu: bool = False) -> np.ndarray:
    """tetra_data (reference layout) of pos's Delaunay tetrahedralisation,
    assembled from x-slabs in parallel (see the module docstring).  margin:
    the slabs' overlap in x (default 8 mean spacings); shell: the depth of
    the y and z boundary layers every slab includes (default 3 spacings).
    gpu: orientation and neighbour slots by the library's GPU assembly
    (sdtfe_tetra_from_simplices) instead of on the host; the same array.
    Raises ValueError when the slabs' tets do not tile the hull."""
    global _POS, _TREE
    from scipy.spatial import ConvexHull, cKDTree
    p = np.ascontiguousarray(np.asarray(pos, dtype=np.float64)[:, :3])
    n = p.shape[0]
    lo3, hi3 = p.min(0), p.max(0)
    spacing = (np.prod(hi3 - lo3) / n) ** (1.0 / 3.0)
    if margin is None:
        margin = 8.0 * spacing
    if shell is None:
        shell = 3.0 * spacing
    near = ((p[:, 1] < lo3[1] + shell) | (p[:, 1] > hi3[1] - shell) |
            (p[:, 2] < lo3[2] + shell) | (p[:, 2] > hi3[2] - shell))
    t0 = time.perf_counter()
    _POS = p
    _TREE = cKDTree(p)
    xs = np.sort(p[:, 0])
    cuts = [xs[0]] + [xs[(n * k) // slabs] for k in range(1, slabs)] + [xs[-1]]
    args = []
    for k in range(slabs):
        lo, hi = cuts[k], cuts[k + 1]
        ext_lo, ext_hi = lo - margin, hi + margin
        idx = np.nonzero(((p[:, 0] >= ext_lo) & (p[:, 0] <= ext_hi)) | near)[0]
        args.append((k, idx, lo, hi, ext_lo, ext_hi, k == 0, k == slabs - 1))
    log(f"[tess] {n} points, {slabs} slabs, margin {margin:.3g}, shell {shell:.3g} "
        f"({int(near.sum())} points), tree {time.perf_counter() - t0:.1f} s")
    parts = [None] * slabs
    ctx = mp.get_context("fork")
    with ctx.Pool(min(workers, slabs)) as pool:
        for k, s, n_unc, n_drop in pool.imap_unordered(_slab, args):
            parts[k] = s
            log(f"[tess] slab {k}: {s.shape[0]} tets ({n_unc} checked against all points, "
                f"{n_drop} dropped), {time.perf_counter() - t0:.1f} s")
    simp = np.concatenate(parts)
    if gpu:
        from .capi import SdtfeError, tetra_from_simplices
        try:
            tetra = tetra_from_simplices(p, simp)
        except SdtfeError as e:
            raise ValueError(f"slab tets do not form a tessellation: {e}") from e
        n_hull = int((tetra[:, 4:] < 0).sum())
    else:
        nb = _neighbors(simp)
        n_hull = int((nb < 0).sum())
    log(f"[tess] neighbours of {simp.shape[0]} tets ({'GPU' if gpu else 'host'}), "
        f"{time.perf_counter() - t0:.1f} s")
    vol = np.abs(circumspheres(p, simp)[2]).sum() / 6.0
    hull = ConvexHull(p)
    log(f"[tess] {simp.shape[0]} tets; volume {vol:.12g} vs hull {hull.volume:.12g}; "
        f"{n_hull} hull faces vs {hull.simplices.shape[0]}; {time.perf_counter() - t0:.1f} s")
    if abs(vol - hull.volume) > 1e-9 * hull.volume or n_hull != hull.simplices.shape[0]:
        raise ValueError("slab tets do not tile the convex hull")
    _POS = _TREE = None
    return tetra if gpu else orient_and_link(p, simp, nb)
