"""Output helpers of the solver (reference: solver/utils.py:21-277).

The reference's `savez` is numpy's uncompressed ``.npz`` writer (a zip of
``.npy`` members named after the keywords); numpy's own writer produces the
same format, so dumps written here are read by reference-era
post-processing with ``numpy.load``, and reference dumps are read here.
Loading never unpickles: object arrays are refused.
"""
import os
import sys

import numpy as np


def savez(file, *args, **kwds):
    """Uncompressed .npz of the given arrays (utils.py:33-112)."""
    np.savez(file, *args, **kwds)


def savez_compressed(file, *args, **kwds):
    """Deflate-compressed .npz (utils.py:114-136)."""
    np.savez_compressed(file, *args, **kwds)


def load(fname):
    """The arrays of a .npz dump as a dict; scalars come back as 0-d
    arrays.  Object (pickled) members raise ValueError."""
    with np.load(fname, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def check_array(x, y):
    """True when the two arrays are equal (utils.py:21-25)."""
    return bool(np.allclose(x, y, atol=1e-16, rtol=0.0))


def mkdir(newdir):
    """mkdir -p (utils.py:242-266)."""
    if os.path.isdir(newdir):
        return
    if os.path.isfile(newdir):
        raise OSError("a file with the same name as the desired dir, '%s', already exists."
                      % newdir)
    os.makedirs(newdir, exist_ok=True)


class PBar:
    """A text progress bar (utils.py:207-240)."""

    def __init__(self, maxval, show=True):
        self.maxval = max(int(maxval), 1)
        self.show = show
        self.count = 0
        self.last = -1

    def update(self):
        self.count += 1
        pct = min(100, int(100 * self.count / self.maxval))
        if self.show and pct != self.last:
            sys.stdout.write("\r[%-50s] %3d%%" % ("#" * (pct // 2), pct))
            sys.stdout.flush()
        self.last = pct

    def finish(self):
        if self.show:
            sys.stdout.write("\n")
            sys.stdout.flush()
