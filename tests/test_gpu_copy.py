"""The two copy kernels behind the measurement figures: swd_calib_copy (its
rocprofv3 readings scale the FD kernels' traffic counters) and swd_copy_peak
(the practical ceiling the roofline reports beside the 8 TB/s specification).
Each call fills the source with a distinct 16-byte pattern per element and
checks the whole destination after the timed copies, so a copy that skipped
elements (a grid or tail mistake) returns an error instead of a rate.  Sizes
cover one element, part of one block's chunk, a ragged tail past one whole
grid stride of each kernel (16384 blocks of 256 or 1024 elements), and a
buffer larger than the Infinity Cache."""
import ctypes

import pytest

from swd_amd import _lib

pytestmark = pytest.mark.gpu

SWD_ERR_ARG = -2
FNS = ("swd_calib_copy", "swd_copy_peak")


def copy_ms(name, nbytes, reps):
    ms = ctypes.c_double(-1.0)
    rc = getattr(_lib.load(), name)(nbytes, reps, ctypes.byref(ms))
    return rc, ms.value


@pytest.mark.parametrize("name", FNS)
@pytest.mark.parametrize("nbytes", [16, 16 * 1023, 16 * 1025, (1 << 20) + 48, 16 * (16384 * 1024 + 37)])
def test_copy_moves_every_element(name, nbytes):
    rc, ms = copy_ms(name, nbytes, 3)
    assert rc == 0, _lib.load().swd_last_error()
    assert ms >= 0


@pytest.mark.parametrize("name", FNS)
def test_copy_rate_of_one_gib_is_below_the_hbm_specification(name):
    """1 GiB each way exceeds the 256 MB Infinity Cache, so read + write
    bytes over the time must stay under the 8 TB/s of HBM3E."""
    nbytes = 1 << 30
    rc, ms = copy_ms(name, nbytes, 5)
    assert rc == 0, _lib.load().swd_last_error()
    tbs = 2 * nbytes / (ms * 1e-3) / 1e12
    assert 1.0 < tbs < 8.0, tbs


@pytest.mark.parametrize("name", FNS)
@pytest.mark.parametrize("nbytes,reps", [(8, 1), (0, 1), (1024, 0)])
def test_copy_rejects_bad_arguments(name, nbytes, reps):
    rc, ms = copy_ms(name, nbytes, reps)
    assert rc == SWD_ERR_ARG
    assert ms == -1.0
