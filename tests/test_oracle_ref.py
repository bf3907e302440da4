"""The oracle's neighbour sets against the reference's own C++ neighbour
search (bench/nnps_c++/nnps.C, built by oracle/Makefile into oracle/_ref/
from the reference sources; skipped where that binary is absent).  Same
strict test d2 < r2 with one radius for every point, in 2D."""
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.skipif(not os.path.exists(O.REF_NNPS2D),
                                reason="reference C++ neighbour search not built")


def _oracle_rows(x, y, radius):
    a = O.make_array(x=x, y=y, h=np.full(len(x), 0.5 * radius))
    off, nb = O.nnps_csr(a, radius_scale=2.0, cell_size=radius)
    return [np.sort(nb[off[i]:off[i + 1]]) for i in range(len(x))]


@pytest.mark.parametrize("kind", ["uniform", "clustered", "lattice"])
def test_oracle_sets_equal_reference_nnps(kind):
    rng = np.random.Generator(np.random.PCG64(11))
    if kind == "uniform":
        x, y = rng.uniform(-1, 1, 3000), rng.uniform(-1, 1, 3000)
        radius = 0.06
    elif kind == "clustered":
        c = rng.uniform(-0.8, 0.8, (12, 2))
        k = rng.integers(0, 12, 3000)
        x, y = c[k, 0] + rng.normal(0, 0.05, 3000), c[k, 1] + rng.normal(0, 0.05, 3000)
        radius = 0.03
    else:
        # a square lattice whose spacing divides the radius: many pairs sit
        # exactly on the cut-off and both sides must exclude them
        i, j = np.meshgrid(np.arange(40), np.arange(40), indexing="ij")
        x, y = i.ravel() * 0.25, j.ravel() * 0.25
        radius = 0.5
    ref = O.ref_nnps2d(x, y, radius)
    got = _oracle_rows(x, y, radius)
    assert sum(len(r) for r in ref) > 5 * len(x)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
