"""Raw float32 file I/O as the reference scripts do it.

* :func:`read_binary_dataset` -- test.py:10-15: float32, C order, reshaped to
  ``(-1, dim[0], dim[1])`` (a stack of slices).
* :func:`read_projections` -- FDK/FDK_test_training.py:9-16: float32, C order,
  ``(num_samples, nview, ncol)``.
* :func:`write_raw` -- test.py:24-26 (``f.write(bytearray(array))``): the raw float32
  bytes of a C-ordered array, no header.
"""
from __future__ import annotations

import numpy as np

__all__ = ["read_binary_dataset", "read_projections", "write_raw"]


def read_binary_dataset(file_name, dim) -> np.ndarray:
    data = np.fromfile(file_name, dtype=np.float32, count=-1)
    return np.reshape(data, (-1, int(dim[0]), int(dim[1])), order="C")


def read_projections(file_name, num_samples, nview, ncol) -> np.ndarray:
    data = np.fromfile(file_name, dtype=np.float32, count=-1)
    return np.reshape(data, (int(num_samples), int(nview), int(ncol)), order="C")


def write_raw(file_name, array) -> None:
    a = np.ascontiguousarray(np.asarray(array, dtype=np.float32))
    with open(file_name, "wb") as f:
        f.write(a.tobytes(order="C"))
