"""The slab-assembled Delaunay tetrahedralisation (host) equals Qhull's over
all points: the same tets, the reference's orientation and neighbour slots."""
import numpy as np
import pytest

from sdtfe_amd import tessellate, workloads
from sdtfe_amd.tessellate_chunked import triangulate_chunked


@pytest.mark.parametrize("kind", ["lattice", "clustered"])
def test_slabs_equal_one_qhull_run(kind):
    if kind == "lattice":
        pos = workloads.jittered_lattice(40, 40, 20, seed=2)
    else:
        pos = workloads.clustered(30000, seed=3)
    got = triangulate_chunked(pos, slabs=5, workers=4, log=lambda m: None)
    ref = tessellate.triangulate(pos)
    assert set(map(tuple, np.sort(got[:, :4], axis=1))) == set(map(tuple, np.sort(ref[:, :4], axis=1)))
    tessellate.check_tetra_data(pos, got)
    v, nb = got[:, :4], got[:, 4:]
    for f in range(4):
        m = nb[:, f] >= 0
        face = np.sort(np.delete(v, f, axis=1)[m], axis=1)
        other = v[nb[m, f]]
        # every vertex of face f is a vertex of the tet across it
        assert (face[:, :, None] == other[:, None, :]).any(axis=2).all()


def test_slabs_that_do_not_tile_raise():
    # a margin of zero leaves tets near the slab cuts unowned: the tiling
    # check must catch it
    pos = workloads.jittered_lattice(24, 24, 12, seed=4)
    with pytest.raises(ValueError):
        triangulate_chunked(pos, slabs=4, workers=2, margin=0.0, shell=0.0, log=lambda m: None)


def test_face_matching_of_int32_simplices():
    """The host face matching packs two vertex ids into one key (id << bits
    | id): with Qhull's int32 simplices over more than 2^16 points that
    overflows 32 bits, so the keys must be built in 64 bits to find
    exactly Qhull's neighbours."""
    from scipy.spatial import Delaunay

    from sdtfe_amd.tessellate import QHULL_OPTIONS
    from sdtfe_amd.tessellate_chunked import _neighbors
    pos = np.random.default_rng(9).random((70000, 3))
    tri = Delaunay(pos, qhull_options=QHULL_OPTIONS)
    assert tri.simplices.dtype == np.int32 and tri.simplices.max() >= 1 << 16
    np.testing.assert_array_equal(_neighbors(tri.simplices), tri.neighbors)
