"""Box arguments of the Python wrappers (host logic, no GPU): with exactly
three frames, per-frame box lengths (3, 3) must stay per-frame lengths for
orthorhombic boxes and not be read as one box matrix; a (3, 3) triclinic
matrix is shared by every frame."""
import numpy as np

from warp_md_amd import _native
from warp_md_amd import autograd, batch

import _gpu

LENS = np.array([[10.0, 11.0, 12.0], [13.0, 14.0, 15.0], [16.0, 17.0, 18.0]], np.float32)


def _diag9(lens):
    out = np.zeros((len(lens), 9), np.float32)
    out[:, 0], out[:, 4], out[:, 8] = lens[:, 0], lens[:, 1], lens[:, 2]
    return out


def test_autograd_three_frames_of_lengths():
    m, vol = autograd._boxes9(LENS, 3, "orthorhombic")
    np.testing.assert_array_equal(m, _diag9(LENS))
    np.testing.assert_allclose(vol, np.prod(LENS.astype(np.float64), axis=1))


def test_autograd_three_frames_triclinic_matrix():
    M = np.array([[10.0, 0, 0], [1.5, 11.0, 0], [0.5, -0.3, 12.0]], np.float32)
    m, vol = autograd._boxes9(M, 3, "triclinic")
    np.testing.assert_array_equal(m, np.tile(M.reshape(1, 9), (3, 1)))
    np.testing.assert_allclose(vol, [10.0 * 11.0 * 12.0] * 3, rtol=1e-12)


def test_batch_three_frames_of_lengths():
    np.testing.assert_array_equal(batch._boxes9(LENS, 3, _native.PBC_ORTHO), _diag9(LENS))
    M = np.array([[10.0, 0, 0], [1.5, 11.0, 0], [0.5, -0.3, 12.0]], np.float32)
    np.testing.assert_array_equal(batch._boxes9(M, 3, _native.PBC_TRICLINIC),
                                  np.tile(M.reshape(1, 9), (3, 1)))


def test_test_helper_three_frames_of_lengths():
    np.testing.assert_array_equal(_gpu.boxes9(LENS, 3, "orthorhombic"), _diag9(LENS))
