"""GPU assembly of the reference's tetra_data from Delaunay simplices
(sdtfe_tetra_from_simplices; what triangulate does with Qhull's facets,
src/triangulate.c:56-175) against the host adapter on the same simplices.
Bar: identical integer arrays (orientation and neighbour slots)."""
import sys
import time

import numpy as np
import pytest
from scipy.spatial import Delaunay

from sdtfe_amd import capi, workloads
from sdtfe_amd.tessellate import QHULL_OPTIONS, orient_and_link
from sdtfe_amd.tessellate_chunked import triangulate_chunked

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("kind,n", [("uniform", 20000), ("clustered", 131072),
                                    ("lattice", 32 * 32 * 16), ("dyadic", 4000)])
def test_assembly_matches_host(kind, n):
    _say(f"[gpu] tet assembly {kind} {n}")
    if kind == "uniform":
        pos = workloads.uniform(n, 3)
    elif kind == "clustered":
        pos = workloads.clustered(n, seed=3)
    elif kind == "lattice":
        pos = workloads.jittered_lattice(32, 32, 16, seed=3)
    else:  # many exactly coplanar and cospherical points
        rng = np.random.default_rng(3)
        pos = np.unique(rng.integers(0, 16, size=(n, 3)).astype(np.float64) / 16.0, axis=0)
    tri = Delaunay(pos, qhull_options=QHULL_OPTIONS)
    want = orient_and_link(pos, tri.simplices, tri.neighbors)
    t0 = time.perf_counter()
    got = capi.tetra_from_simplices(pos, tri.simplices)
    _say(f"  {want.shape[0]} tets in {1e3 * (time.perf_counter() - t0):.1f} ms")
    np.testing.assert_array_equal(got, want)


def test_slab_tessellation_on_gpu_matches_host():
    _say("[gpu] slab tessellation, GPU assembly vs host")
    pos = workloads.jittered_lattice(40, 40, 20, seed=2)
    host = triangulate_chunked(pos, slabs=4, workers=4, log=lambda m: None)
    gpu = triangulate_chunked(pos, slabs=4, workers=4, log=lambda m: None, gpu=True)
    np.testing.assert_array_equal(gpu, host)


def test_assembly_errors():
    _say("[gpu] tet assembly errors")
    pos = workloads.uniform(500, 1)
    simp = Delaunay(pos, qhull_options=QHULL_OPTIONS).simplices
    with pytest.raises(capi.SdtfeError, match="more than two"):
        capi.tetra_from_simplices(pos, np.concatenate([simp, simp[:5]]))
    bad = simp.copy()
    bad[7, 2] = 500
    with pytest.raises(capi.SdtfeError, match="out of range"):
        capi.tetra_from_simplices(pos, bad)
    assert capi.tetra_from_simplices(pos, simp[:0]).shape == (0, 8)
