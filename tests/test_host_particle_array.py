"""ParticleArray's tag semantics on the host, restating the expectations of
the reference's own tests (base/tests/particle_array_tests.py): LocalReal
particles first in the reference's alignment order, num_real_particles,
get(only_real_particles=...), set with tags and short arrays, removal by
swap-with-last, append/extend/add/copy, and pickling."""
import pickle

import numpy as np
import pytest

from pysph_amd.base.particle_array import ParticleArray, align_index, get_particle_array
from pysph_amd.base.particle_types import LocalDummy, LocalReal, RemoteReal


def _loop_align(tag):
    """particle_array.pyx:954-1040 step by step (the swap walk)."""
    n = len(tag)
    idx = [None] * n
    nxt = 0
    for i in range(n):
        if tag[i] == LocalReal:
            if i != nxt:
                idx[nxt], idx[i] = i, idx[nxt]
            else:
                idx[i] = i
            nxt += 1
        else:
            idx[i] = i
    return np.array(idx, dtype=np.int64), nxt


def test_align_index_matches_the_swap_walk():
    rng = np.random.Generator(np.random.PCG64(3))
    for _ in range(2000):
        n = int(rng.integers(0, 50))
        tag = np.where(rng.random(n) < rng.random(), LocalReal, rng.integers(1, 5, n))
        a, m = align_index(tag)
        b, mb = _loop_align(list(tag))
        assert m == mb and np.array_equal(a, b)


def test_constructor_with_tags_aligns():
    # particle_array_tests.py:82-94
    p = ParticleArray(x={"data": [1, 2, 3, 4.]}, y={"data": [0, 1, 2, 3.]},
                      z={"data": [0, 0, 0, 0.]}, tag={"data": [0, 1, 0, 1], "type": "long"})
    assert list(p.get("tag", only_real_particles=False)) == [0, 0, 1, 1]
    assert list(p.get("x", only_real_particles=False)) == [1, 3, 2, 4]
    assert list(p.get("y", only_real_particles=False)) == [0, 2, 1, 3]
    assert p.num_real_particles == 2
    assert list(p.get("x")) == [1, 3] and list(p.x) == [1, 3]


def test_defaults_without_data():
    # particle_array_tests.py:97-114
    p = ParticleArray(x={}, y={}, z={}, h={})
    assert p.get_number_of_particles() == 0 and p.has_array("tag")
    p = ParticleArray(x={"default": 10.0}, y={"data": [1.0, 2.0]}, z={},
                      h={"data": [0.1, 0.1]})
    assert p.get_number_of_particles() == 2
    assert list(p.x) == [10., 10.] and list(p.z) == [0, 0]


def test_set_tags_short_and_long_arrays():
    # particle_array_tests.py:149-179
    p = ParticleArray(x={"data": [1, 2, 3, 4.]}, y={"data": [0, 1, 2, 3.]},
                      z={"data": [0] * 4}, m={"data": [1.] * 4}, h={"data": [.1] * 4})
    p.set(x=[4., 3, 2, 1], h=[0.2] * 4)
    assert list(p.get("x")) == [4., 3, 2, 1] and list(p.get("h")) == [0.2] * 4
    p.set(tag=[0, 1, 1, 1])
    assert list(p.get("tag", only_real_particles=False)) == [0, 1, 1, 1]
    assert list(p.get("tag")) == [0]
    p.set(x=[5, 6, 7])
    assert list(p.get("x", only_real_particles=False)) == [5, 6, 7, 1]
    with pytest.raises(ValueError):
        p.set(x=[1., 2, 3, 5, 6])


def test_align_particles_order_and_dirty_flag():
    # particle_array_tests.py:490-514
    p = ParticleArray()
    p.add_property({"name": "x", "data": [1, 2, 3, 4, 5, 6, 7, 8, 9, 10]})
    p.add_property({"name": "y", "data": [10, 9, 8, 7, 6, 5, 4, 3, 2, 1]})
    p.set_dirty(False)
    p.set(tag=[0, 0, 1, 1, 1, 0, 4, 0, 1, 5])
    assert list(p.get("x", only_real_particles=False)) == [1, 2, 6, 8, 5, 3, 7, 4, 9, 10]
    assert list(p.get("y", only_real_particles=False)) == [10, 9, 5, 3, 6, 8, 4, 7, 2, 1]
    assert p.is_dirty
    p.set_dirty(False)
    p.set(tag=[0, 0, 0, 0, 1, 1, 1, 1, 1, 1])
    assert list(p.get("x", only_real_particles=False)) == [1, 2, 6, 8, 5, 3, 7, 4, 9, 10]
    assert not p.is_dirty
    assert p.num_real_particles == 4


def test_remove_particles_swaps_in_the_last():
    # particle_array_tests.py:279-325
    p = ParticleArray(x={"data": [1, 2, 3, 4.]}, y={"data": [0, 1, 2, 3.]},
                      z={"data": [0] * 4}, m={"data": [1.] * 4}, h={"data": [.1] * 4})
    p.add_temporary_array("tmp1")
    p.remove_particles([0, 1])
    assert p.get_number_of_particles() == 2
    assert list(p.x) == [3., 4.] and list(p.y) == [2., 3.] and len(p.tmp1) == 2
    with pytest.raises(ValueError):
        p.remove_particles(list(range(10)))
    p.remove_particles([2])
    assert list(p.x) == [3., 4.]


def test_remove_tagged_particles():
    # particle_array_tests.py:387-427
    p = ParticleArray(x={"data": [1, 2, 3, 4.]}, y={"data": [0, 1, 2, 3.]},
                      z={"data": [0] * 4}, m={"data": [1.] * 4}, h={"data": [.1] * 4},
                      tag={"data": [1, 1, 1, 0]})
    p.add_temporary_array("tmp1")
    p.remove_tagged_particles(0)
    assert p.get_number_of_particles() == 3
    assert list(p.get("x", only_real_particles=False)) == [1, 2, 3.]
    assert list(p.get("y", only_real_particles=False)) == [0., 1, 2]
    assert len(p.get("tmp1", only_real_particles=False)) == 3
    assert len(p.x) == 0 and len(p.tmp1) == 0 and p.num_real_particles == 0


def test_extend_append_copy():
    # particle_array_tests.py:474-488
    p = ParticleArray(default_particle_tag=10, x={}, y={"default": -1.})
    p.extend(5)
    assert list(p.get("x", only_real_particles=False)) == [0] * 5
    assert list(p.get("y", only_real_particles=False)) == [-1.] * 5
    assert list(p.get("tag", only_real_particles=False)) == [10] * 5
    # particle_array_tests.py:516-538
    p1 = ParticleArray()
    p1.add_property({"name": "x", "data": [1, 2, 3]})
    p1.align_particles()
    p2 = ParticleArray(x={"data": [4, 5, 6]}, y={"data": [1, 2, 3]}, tag={"data": [1, 0, 1]})
    p1.append_parray(p2)
    assert p1.get_number_of_particles() == 6
    assert list(p1.x) == [1, 2, 3, 5] and list(p1.y) == [0, 0, 0, 2]
    assert list(p1.tag) == [0, 0, 0, 0]
    # particle_array_tests.py:540-563
    p1 = ParticleArray()
    p1.add_property({"name": "x", "data": list(range(1, 11))})
    p1.add_property({"name": "y"})
    p1.add_property({"name": "t"})
    p1.align_particles()
    p2 = ParticleArray()
    p2.add_property({"name": "t", "data": [-1, -1, -1, -1]})
    p2.add_property({"name": "s", "data": [2, 3, 4, 5]})
    p2.align_particles()
    p1.copy_properties(p2, start_index=5, end_index=9)
    assert list(p1.t) == [0, 0, 0, 0, 0, -1, -1, -1, -1, 0]
    p1.add_property({"name": "s"})
    p1.copy_properties(p2, start_index=5, end_index=9)
    assert list(p1.s) == [0, 0, 0, 0, 0, 2, 3, 4, 5, 0]


def test_add_particles_and_extract_align():
    p = get_particle_array(x=[0.0, 1.0, 2.0], name="a")
    p.add_particles(x=np.array([3.0, 4.0]), tag=np.array([RemoteReal, LocalReal]))
    assert p.get_number_of_particles() == 5 and p.num_real_particles == 4
    assert list(p.get("tag", only_real_particles=False)) == [0, 0, 0, 0, RemoteReal]
    assert list(p.x) == [0.0, 1.0, 2.0, 4.0]
    q = p.extract_particles([4, 1])
    assert list(q.get("x", only_real_particles=False)) == [1.0, 3.0]
    assert q.num_real_particles == 1


def test_pickle_round_trip():
    # particle_array_tests.py:565-578
    p1 = ParticleArray(name="fluid", cl_precision="single")
    p1.add_property({"name": "x", "data": [1, 2, 3, 4, 5, 6, 7, 8, 9, 10]})
    p1.add_property({"name": "y"})
    p1.add_property({"name": "t", "default": 2.5})
    p1.set(tag=[0, 1, 0, 0, 0, 2, 0, 0, 1, 0])
    p1.add_temporary_array("_tmp")
    p1.constants["c"] = 3.0
    p2 = pickle.loads(pickle.dumps(p1))
    assert p2.name == "fluid" and p2.dtype == p1.dtype
    assert len(p1.x) == len(p2.x) == 7 and np.array_equal(p1.x, p2.x)
    for k in ("x", "y", "t", "tag"):
        assert np.array_equal(p1.get(k, only_real_particles=False),
                              p2.get(k, only_real_particles=False)), k
    assert p2.num_real_particles == 7 and p2.c == 3.0
    assert p2.default_values["t"] == 2.5
    assert "_tmp" in p2.temporary_arrays and "_tmp" not in p2.properties
