"""Host pre-processing of the pydtfe mirror against the reference's
convert_to_qhdata / rotate3d (golden vectors) and the binding's errors."""
import numpy as np
import pytest

from sdtfe_amd import pydtfe


def test_convert_and_rotate_match_reference(golden):
    g = golden("g5_utils.npz")
    qh = pydtfe.convert_to_qhdata(np.asarray(g["pos32"], order="F"))
    np.testing.assert_array_equal(qh[:, :3], g["qh"])
    pydtfe.rotate3d(qh, qh.shape[0], g["theta"], g["center"])
    np.testing.assert_array_equal(qh[:, :3], g["rotated"])


def test_shuffle_is_a_partial_permutation():
    data = np.arange(40, dtype=np.float64).reshape(10, 4)
    d = data.copy()
    pydtfe.shuffle_buff(d, 10, 6, seed=3)
    assert sorted(d[:, 0].tolist()) == sorted(data[:, 0].tolist())
    assert (d[:, 1] - d[:, 0] == 1).all()  # rows move whole
    d2 = data.copy()
    pydtfe.shuffle_buff(d2, 10, 6, seed=3)
    np.testing.assert_array_equal(d, d2)


def test_binding_errors():
    with pytest.raises(ValueError, match="incorrect shape"):
        pydtfe.compute_surface_density(np.zeros((10, 2)), [0, 0, 0], 8, 1.0, 1, 1, 1, 0, 1)
    with pytest.raises(ValueError, match="not enough particles"):
        pydtfe.compute_surface_density(np.zeros((10, 3)), [0, 0, 0], 8, 0.3, 1, 1, 1, 0, 1)
