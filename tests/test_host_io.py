"""Output formats on the host (no GPU): the .npz writer/reader of
solver/utils.py and the cell-id rule of the neighbour dumps."""
import math

import numpy as np
import pytest

from pysph_amd.base.cell import real_to_int
from pysph_amd.solver import utils
from pysph_amd.solver.post_step_functions import PrintNeighborInformation, cell_ids


def test_savez_round_trip(tmp_path):
    x = np.linspace(0, 1, 7)
    f = str(tmp_path / "solution_fluid_0.1.npz")
    utils.savez(f, dt=1e-4, cell_size=0.02, np=7, x=x, rho=np.full(7, 1000.0))
    d = utils.load(f)
    assert set(d) == {"dt", "cell_size", "np", "x", "rho"}
    assert float(d["dt"]) == 1e-4 and int(d["np"]) == 7
    assert np.array_equal(d["x"], x)
    # the reference reads these with numpy.load
    with np.load(f) as z:
        assert np.array_equal(z["rho"], np.full(7, 1000.0))


def test_savez_compressed_and_unnamed(tmp_path):
    f = str(tmp_path / "c")
    utils.savez_compressed(f, np.arange(5), y=np.ones(3))
    d = utils.load(f + ".npz")
    assert np.array_equal(d["arr_0"], np.arange(5)) and np.array_equal(d["y"], np.ones(3))


def test_load_refuses_object_members(tmp_path):
    f = str(tmp_path / "obj.npz")
    np.savez(f, a=np.array([{"k": 1}], dtype=object))
    with pytest.raises(ValueError):
        utils.load(f)


def test_cell_ids_follow_real_to_int():
    rng = np.random.Generator(np.random.PCG64(0))
    x, y, z = (rng.normal(size=200) for _ in range(3))
    c = cell_ids(x, y, z, 0.13)
    for k in range(200):
        assert tuple(c[k]) == (real_to_int(x[k], 0.13), real_to_int(y[k], 0.13),
                               real_to_int(z[k], 0.13))
    assert c.dtype == np.int32
    assert tuple(cell_ids([-1e-12], [0.0], [0.26], 0.13)[0]) == (-1, 0, int(math.floor(2.0)))


def test_print_neighbor_information_rejects_unknown_format():
    with pytest.raises(ValueError):
        PrintNeighborInformation(fmt="hdf5")


def test_mkdir_and_pbar(tmp_path, capsys):
    d = tmp_path / "a" / "b"
    utils.mkdir(str(d))
    utils.mkdir(str(d))
    assert d.is_dir()
    bar = utils.PBar(4)
    for _ in range(4):
        bar.update()
    bar.finish()
    assert "100%" in capsys.readouterr().out
