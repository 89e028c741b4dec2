"""The BBKS table writer: the published fit, the CLASS normalisation of the
reference's own table, and the potential generateIC_basic rebuilds from it."""
import os
import sys

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "gevolution-1.2_amd"))
import bbks  # noqa: E402

GOLD = os.path.join(HERE, "golden")


def test_bbks_fit_limits_and_a_value():
    assert bbks.bbks_T(1e-9, 0.31, 0.67) == pytest.approx(1.0, abs=1e-7)
    # q = 1: ln(3.34)/2.34 * (1 + 3.89 + 16.1^2 + 5.46^3 + 6.71^4)^-1/4
    want = np.log(3.34) / 2.34 * (1 + 3.89 + 16.1 ** 2 + 5.46 ** 3 + 6.71 ** 4) ** -0.25
    assert bbks.bbks_T(0.31 * 0.67, 0.31, 0.67) == pytest.approx(want, rel=1e-14)
    k = np.logspace(-4, 2, 50)
    assert np.all(np.diff(bbks.bbks_T(k, 0.31, 0.67)) < 0)


def test_superhorizon_normalisation_matches_reference_class_table():
    """On superhorizon scales d_tot -> -2 Phi = -6/5 and theta = (2/3) k^2 Phi / Hc;
    the reference's CLASS table at z = 100 agrees to a few per cent (its
    background has radiation, 3% of the matter density at z = 100, which
    the matter-era table and the simulated background leave out)."""
    k_c, d_c, t_c = O.load_transfer_functions(os.path.join(GOLD, "class_tk.dat"), "tot", 1.0, 1.0)
    sel = k_c < 3e-4
    k, d, t = bbks.bbks_table(k_c[sel], 100.0, 0.31, 0.67556)
    assert np.max(np.abs(d / d_c[sel] - 1)) < 0.03
    assert np.max(np.abs(t / t_c[sel] - 1)) < 0.06


def test_ic_potential_is_minus_k2_phi(tmp_path):
    """generateIC_basic's phi ordinate (1.5 Hc^2 delta + rescale theta/k^2) w
    reduces to -k^2 Phi w for the matter-era table (box units)."""
    boxsize, h, Om = 2048.0, 0.67, 0.31
    path = str(tmp_path / "bbks_tk.dat")
    bbks.write_class_tk(path, 100.0, Om, h)
    cosmo = O.make_cosmo(0.26, 0.05, h)
    fpg = 1.5 * boxsize ** 2 / 2997.92458 ** 2
    sim = O.Sim(64, 1, 0, 48.0, 0.04, 30.0, fpg, cosmo)
    ic = O.IC(7, 1, 1, 16, 1, 100.0, 2.1e-9, 0.965, 0.05, boxsize)
    k, dt, tt = O.load_transfer_functions(path, "tot", boxsize, h)
    _, dc, tc = O.load_transfer_functions(path, "cdm", boxsize, h)
    yp, yd, yv = O.ic_tables(sim, ic, k, dt, tt, dc, tc)
    w = np.pi * np.sqrt(2.1e-9 * (k * h / boxsize / 0.05) ** (0.965 - 1.0) / k) / k
    Phi = 0.6 * bbks.bbks_T(k / boxsize, Om, h)
    # Lambda at z = 100 perturbs the matter-era identities at the 1e-6 level
    assert np.max(np.abs(yp / (-(k ** 2) * Phi * w) - 1.0)) < 1e-4
