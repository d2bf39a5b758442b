"""The streaming calibration kernel on the GPU: both load forms report a
rate, in GB/s of bytes read plus written, of the order of the HBM bandwidth
(the bench's practical ceiling beside the datasheet peak)."""
import ctypes as C

import pytest

import dgk

pytestmark = pytest.mark.gpu


def test_calibrate_stream_rates():
    a, r = C.c_double(0.0), C.c_double(0.0)
    assert dgk.lib().dgk_calibrate_stream(1 << 22, 3, C.byref(a), C.byref(r)) == dgk.OK
    # five arrays of 32 MiB read and one written, beyond the L2: a working
    # kernel streams at more than 1 TB/s and below the memory-side cache's rate
    assert 1000.0 < a.value < 20000.0
    assert 1000.0 < r.value < 20000.0
