"""Host-side argument checks of the Python binding (scft_amd.Context): arrays
of the wrong length are refused before any call into the library, which
reads M complex values (or 3 per plate) from each pointer it is given."""
import numpy as np
import pytest

import scft_amd
from scft_amd import Workload


def _unopened_context(**w):
    # the checks need only the parameters; no device context is created
    c = object.__new__(scft_amd.Context)
    c.params = [Workload(**w).params()]
    c.h = None
    return c


def test_wrong_sizes_are_refused_before_the_library():
    c = _unopened_context(grid=(8, 8, 8), n_plates=2)
    M = c.M()
    assert M == 512
    with pytest.raises(ValueError, match="set_fields"):
        c.set_fields(np.zeros(M), np.zeros(M), np.zeros(M - 1))
    with pytest.raises(ValueError, match="fft"):
        c.fft(np.zeros(M + 1), -1)
    with pytest.raises(ValueError, match="centres"):
        c.set_plates([0.5, 0.5, 0.5], [[0, 0, 1], [0, 0, 1]])
    with pytest.raises(ValueError, match="orientations"):
        c.set_plates([[0.4, 0.5, 0.5], [0.6, 0.5, 0.5]], [0, 0, 1])
    with pytest.raises(ValueError, match="axis"):
        c.grad_plates(axis=(0.0, 1.0))
    with pytest.raises(ValueError, match="wabm"):
        c.set_array("wabm", np.zeros(M // 2))


def test_one_plate_system_takes_three_values_per_array():
    c = _unopened_context(grid=(8, 8, 8), n_plates=1)
    with pytest.raises(ValueError, match="centres"):
        c.set_plates([[0.4, 0.5, 0.5], [0.6, 0.5, 0.5]], [0, 0, 1])
