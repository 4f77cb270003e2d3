"""The list entry points' bridge (wgfft/_rows.c) on the CPU: which rows it
passes to wgfft_exec_host_rows where they lie, and which it sends back for
conversion.  With a NULL handle the library rejects the call before any
device work, so the bridge is exercised without a GPU."""
import numpy as np
import pytest

import wgfft
from wgfft import _lib


def _call(rows, n, h=0):
    bridge, fn = _lib.rows_bridge()
    out = np.empty((8, n), dtype=np.complex64)
    return bridge.exec_rows(fn, h, rows, out.ctypes.data, n, wgfft.WGFFT_FORWARD)


def test_bridge_loads_and_points_at_the_entry():
    bridge, fn = _lib.rows_bridge()
    assert callable(bridge.exec_rows) and fn


@pytest.mark.parametrize("rows", [
    [np.zeros(8, np.complex64)] * 3,
    list(np.zeros((4, 8), np.complex64)),
    tuple(np.ones((2, 8), np.complex64)),
    np.zeros((5, 8), np.complex64),            # a 2-D array is a sequence of rows
    [np.zeros(8, np.complex64).view(np.complex64)],
])
def test_contiguous_complex64_rows_reach_the_library(rows):
    rc, bad = _call(rows, 8)
    assert bad == -1
    assert rc == wgfft.WGFFT_ERR_INVALID_ARGUMENT and "handle" in _lib.last_error()


@pytest.mark.parametrize("rows,bad", [
    ([np.zeros(8, np.complex64), np.zeros(8, np.complex128)], 1),   # complex128
    ([np.zeros(16, np.complex64)[::2]], 0),                          # strided
    ([np.zeros(8, np.complex64), [0j] * 8], 1),                      # a Python list
    ([np.zeros(4, np.complex64)], 0),                                # wrong length
    ([np.zeros(16, np.float32)], 0),                                 # interleaved floats, not complex
    ([np.zeros(8, np.complex64).astype(">c8")], 0),                  # byte-swapped
])
def test_other_rows_are_sent_back_for_conversion(rows, bad):
    assert _call(rows, 8) == (0, bad)


def test_not_a_sequence_raises():
    with pytest.raises(TypeError):
        _call(5, 8)
