"""Host-only entry points of the IC generator (no device work): the template
and transfer-function loaders of the HIP library against the reference's
fixtures and the oracle's independent parsers.  CPU."""
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name,n", [("sc0", 64), ("sc1", 64), ("bcc", 16), ("fcc", 32)])
def test_template_loader_matches_fixture(gev, name, n):
    path = os.path.join(GOLD, name + "_crystal.dat")
    t = gev.load_homogeneous_template(path)
    assert t.shape == (n, 3) and t.dtype == np.float32
    assert np.array_equal(t, O.load_template(path))
    assert t.min() >= 0.0 and t.max() < 1.0


def test_sc1_template_is_the_cell_centred_lattice(gev):
    t = gev.load_homogeneous_template(os.path.join(GOLD, "sc1_crystal.dat"))
    want = set(((i + 0.5) / 4, (j + 0.5) / 4, (k + 0.5) / 4) for i in range(4) for j in range(4) for k in range(4))
    assert set(map(tuple, t.astype(np.float64).tolist())) == want


@pytest.mark.parametrize("q", ["tot", "cdm", "b"])
def test_transfer_loader_matches_parser(gev, q):
    path = os.path.join(GOLD, "class_tk.dat")
    k, d, t = gev.load_transfer_functions(path, q, 320.0, 0.67556)
    ko, do, to = O.load_transfer_functions(path, q, 320.0, 0.67556)
    assert len(k) == 135
    assert np.array_equal(k, ko) and np.array_equal(d, do) and np.array_equal(t, to)
    assert np.all(np.diff(k) > 0)


def test_loader_errors(gev, tmp_path):
    with pytest.raises(gev.GevError):
        gev.load_transfer_functions(str(tmp_path / "missing.dat"), "tot", 1.0, 1.0)
    with pytest.raises(gev.GevError):
        gev.load_transfer_functions(os.path.join(GOLD, "class_tk.dat"), "nosuch", 1.0, 1.0)
    bad = tmp_path / "unordered.dat"
    bad.write_text("# 1:k 2:d_tot 3:t_tot\n1.0 2.0 3.0\n0.5 2.0 3.0\n")
    with pytest.raises(gev.GevError):
        gev.load_transfer_functions(str(bad), "tot", 1.0, 1.0)
    with pytest.raises(gev.GevError):
        gev.load_homogeneous_template(os.path.join(GOLD, "class_tk.dat"))
