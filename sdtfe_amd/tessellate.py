"""Host tessellation adapter: Qhull Delaunay -> the reference ``tetra_data`` layout.

The reference builds its tetrahedra with Qhull (``triangulate``,
src/triangulate.c:8-181, options ``"qhull d Qz Qt Qbb"``, src/main.c:160) and
then fixes two conventions the walk depends on:

* orientation (src/triangulate.c:72-99): B and C are swapped when
  ``(AB x AC) . AD > 0``, so every tet has ``det(AB, AC, AD) <= 0``;
* neighbour slots (src/triangulate.c:111-175): row slot ``4 + f`` holds the
  tet across the face opposite vertex ``f`` (faces BDC, ACD, ADB, ABC), ``-1``
  on the hull.

Qhull itself stays on the host.  The only Qhull in this image is the copy
bundled with scipy; its ``simplices`` list the vertices in the same
``facet->vertices`` order and its tets in the same facet order the reference
reads, and its ``neighbors[:, k]`` is the tet opposite vertex ``k``.  The
orientation test below uses the reference's expression order so the swap
decision is bit-identical for the same vertex order.
"""
from __future__ import annotations

import numpy as np

QHULL_OPTIONS = "Qz Qt Qbb"  # scipy adds "d"; matches src/main.c:160


def orient_and_link(pos: np.ndarray, simplices: np.ndarray,
                    neighbors: np.ndarray) -> np.ndarray:
    """Return ``tetra_data`` (T x 8 int32) from Qhull simplices/neighbours.

    ``pos`` is (N, >=3) float64.  Rows are (A, B, C, D, N_BDC, N_ACD, N_ADB,
    N_ABC).
    """
    simp = np.ascontiguousarray(simplices, dtype=np.int64)
    nb = np.ascontiguousarray(neighbors, dtype=np.int64)
    p = np.asarray(pos, dtype=np.float64)
    a, b, c, d = (p[simp[:, k], :3] for k in range(4))
    ab = b - a
    ac = c - a
    ad = d - a
    # norm = ab x ac, then the dot with ad summed from 0 (src/triangulate.c:80-90)
    n0 = ab[:, 1] * ac[:, 2] - ab[:, 2] * ac[:, 1]
    n1 = ab[:, 2] * ac[:, 0] - ab[:, 0] * ac[:, 2]
    n2 = ab[:, 0] * ac[:, 1] - ab[:, 1] * ac[:, 0]
    temp = 0.0 + n0 * ad[:, 0]
    temp = temp + n1 * ad[:, 1]
    temp = temp + n2 * ad[:, 2]
    swap = temp > 0
    out = np.empty((simp.shape[0], 8), dtype=np.int32)
    out[:, 0:4] = simp
    out[:, 4:8] = nb
    # swapping vertices B and C also swaps the faces opposite them
    out[swap, 1], out[swap, 2] = simp[swap, 2], simp[swap, 1]
    out[swap, 5], out[swap, 6] = nb[swap, 2], nb[swap, 1]
    return out


def triangulate(pos: np.ndarray, gpu: bool = False) -> np.ndarray:
    """Delaunay-tetrahedralise ``pos`` (N x 3 or N x 4 float64) with Qhull on
    the host.

    Returns ``tetra_data`` in the reference layout.  Mirrors ``triangulate``
    (src/triangulate.c:8-181).  gpu: orientation and neighbour slots are
    assembled on the GPU (sdtfe_tetra_from_simplices; the same array)
    instead of from Qhull's neighbours on the host.  Raises ``ValueError``
    for fewer than 4 points or a degenerate (flat) point set, where Qhull
    fails.
    """
    from scipy.spatial import Delaunay  # host dependency, imported lazily

    p = np.ascontiguousarray(np.asarray(pos, dtype=np.float64)[:, :3])
    if p.shape[0] < 4:
        raise ValueError("need at least 4 points to tetrahedralise")
    try:
        tri = Delaunay(p, qhull_options=QHULL_OPTIONS)
    except Exception as exc:  # scipy raises QhullError
        raise ValueError(f"Qhull failed: {exc}") from exc
    if gpu:
        from .capi import tetra_from_simplices
        return tetra_from_simplices(p, tri.simplices)
    return orient_and_link(p, tri.simplices, tri.neighbors)


def to_particle_data(pos: np.ndarray) -> np.ndarray:
    """(N, 3) positions -> the reference's N x 4 AoS float64 buffer
    (x, y, z, w) with w = 0 (``convert_to_qhdata``, src/utils.c:8-16)."""
    p = np.asarray(pos, dtype=np.float64)
    out = np.zeros((p.shape[0], 4), dtype=np.float64)
    out[:, :3] = p[:, :3]
    return out


def check_tetra_data(pos: np.ndarray, tetra: np.ndarray) -> None:
    """Assert the layout invariants the walk relies on (tests and callers
    that bring their own tessellation)."""
    t = np.asarray(tetra)
    n_t = t.shape[0]
    if t.ndim != 2 or t.shape[1] != 8:
        raise ValueError("tetra_data must be T x 8")
    if n_t and (t[:, :4].min() < 0 or t[:, :4].max() >= pos.shape[0]):
        raise ValueError("vertex index out of range")
    nbr = t[:, 4:]
    if n_t and (nbr.min() < -1 or nbr.max() >= n_t):
        raise ValueError("neighbour index out of range")
