"""Host-side typed arrays returned by the locator API.

The reference's LongArray/DoubleArray (carray_pyx.src:189-330) are growable
C arrays with numpy views.  Here they only carry query results back to the
host (neighbour indices), so they are thin wrappers over numpy arrays with
the reference's accessors.
"""
import numpy as np


class _BaseArray:
    _dtype = np.float64

    def __init__(self, n=0):
        if isinstance(n, (int, np.integer)):
            self._data = np.zeros(int(n), dtype=self._dtype)
        else:
            self._data = np.ascontiguousarray(n, dtype=self._dtype)

    @property
    def length(self):
        return len(self._data)

    def __len__(self):
        return len(self._data)

    def __getitem__(self, i):
        return self._data[i]

    def __setitem__(self, i, v):
        self._data[i] = v

    def get_npy_array(self):
        return self._data

    def set_data(self, arr):
        self._data = np.ascontiguousarray(arr, dtype=self._dtype)

    def append(self, v):
        self._data = np.append(self._data, self._dtype(v))

    def reset(self):
        self._data = self._data[:0]

    def resize(self, n):
        d = np.zeros(n, dtype=self._dtype)
        k = min(n, len(self._data))
        d[:k] = self._data[:k]
        self._data = d

    def __repr__(self):
        return "%s(%r)" % (type(self).__name__, self._data)


class LongArray(_BaseArray):
    _dtype = np.int64


class IntArray(_BaseArray):
    _dtype = np.int32


class DoubleArray(_BaseArray):
    _dtype = np.float64


class FloatArray(_BaseArray):
    _dtype = np.float32
