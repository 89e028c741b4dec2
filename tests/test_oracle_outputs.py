"""Pins the oracle's main loop with power-spectrum outputs (main.cpp:605-693,
EXACT_OUTPUT_REDSHIFTS; writePowerSpectrum, tools.hpp:268-346) against a
second oracle box stepped by hand: the output is the linear interpolation in
redshift (numpy.interp) between the delta spectra of the two cycles around
z_pk, the loop stops after the field stage of the crossing cycle, and the
split cycle (field stage + particle stage) is the whole cycle.  CPU only."""
import ctypes

import numpy as np
import pytest

import oracle as O
from states import gr_state, sim_params


def _boxes(N, seed, gr):
    st = gr_state(N, seed)
    sp = sim_params(N, 5.0 * N, gr=gr)
    cosmo = O.make_cosmo()
    sim = O.Sim(N, sp["gr_flag"], 0, sp["Cf"], sp["steplimit"], sp["z_relax"], sp["fourpiG"], cosmo)
    a0 = 1.0 / (1.0 + sp["z_in"])
    dt0 = O.lib().orc_next_dtau(ctypes.byref(sim), a0)
    boxes = []
    for _ in range(2):
        ob = O.OracleBox(N, len(st["ids"]))
        ob.pos[:] = st["pos"]; ob.q[:] = st["q"]; ob.id[:] = st["ids"]; ob.box.pmass = st["pmass"]
        ob.phi[:] = st["phi"]; ob.chi[:] = st["chi"]; ob.Bi[:] = st["Bi"]
        ob.BiFT[:] = st["BiFT"]; ob.scalarFT[:] = st["scalarFT"]
        ob.file()
        boxes.append(ob)
    states = [O.State(a0, 0.0, dt0, 0.0, 0.0, 0.0, 0.0, 0) for _ in range(2)]
    return sim, boxes, states


def _delta_power(ob, sim, st, numbins):
    """The delta spectrum of output.hpp by direct oracle calls (not the loop)."""
    N = ob.N
    ob.source[:] = 0.0
    if sim.gr_flag:
        O.projection_T00_project(N, ob.pos, ob.q, ob.id, ob.box.pmass, st.a, ob.phi, 1.0, out=ob.source)
    else:
        O.scalar_projection_cic_project(N, ob.pos, ob.id, ob.box.pmass, 1.0, out=ob.source)
    F = O.fft_r2c(ob.source)
    return O.extract_cross_spectrum(F, None, numbins, 1, 1)


def test_split_cycle_is_the_whole_cycle():
    sim, (a, b), (sa, sb) = _boxes(16, 3, True)
    for _ in range(3):
        a.cycle(sim, sa)
        b.cycle_fields(sim, sb)
        b.cycle_particles(sim, sb)
    for name in ("pos", "q", "phi", "chi", "Bi", "BiFT"):
        assert np.array_equal(getattr(a, name), getattr(b, name)), name
    assert (sa.a, sa.tau, sa.dtau, sa.cycle) == (sb.a, sb.tau, sb.dtau, sb.cycle)
    a.close(); b.close()


@pytest.mark.parametrize("gr", [True, False])
def test_output_is_the_redshift_interpolation_of_the_two_cycles_around_z(gr):
    N, numbins = 16, 8
    sim, (a, b), (sa, sb) = _boxes(N, 5, gr)
    z_target = 96.0  # a few cycles after z_in = 100
    boxsize = 5.0 * N
    out = a.run_to_outputs(sim, sa, [z_target], numbins, boxsize)
    c_out = int(out["cycle"][0])
    assert sa.cycle == c_out and out["z"][0] == 0.0 + (1.0 / (1.0 / (z_target + 1.0)) - 1.0)
    # step the second box by hand: the spectra of cycles c_out - 1 and c_out
    for c in range(c_out):
        b.cycle_fields(sim, sb)
        if c == c_out - 1:
            z1 = 1.0 / sb.a - 1.0
            k1, p1, _, _, occ1 = _delta_power(b, sim, sb, numbins)
        b.cycle_particles(sim, sb)
    b.cycle_fields(sim, sb)
    z2 = 1.0 / sb.a - 1.0
    k2, p2, _, _, occ2 = _delta_power(b, sim, sb, numbins)
    assert z1 >= z_target > z2, (z1, z_target, z2)
    # the crossing cycle is the first with 1/a < z + 1: the loop stopped there
    assert np.array_equal(a.pos, b.pos) and np.array_equal(a.q, b.q) and np.array_equal(a.phi, b.phi)
    om = 0.26 + 0.05 if gr else O.make_cosmo().Omega_m
    rescalep = float(N) ** 6 * 2.0 * np.pi ** 2 * om * om
    live = occ2 > 0
    assert np.array_equal(occ1, occ2) and np.array_equal(out["occupation"][0], occ2)
    expect = np.array([np.interp(z_target, [z2, z1], [p2[i], p1[i]]) for i in range(numbins)]) / rescalep
    # the look-ahead spectrum passes through the text table ("%e": 7 digits)
    np.testing.assert_allclose(out["pk"][0][live], expect[live], rtol=1e-6, atol=0)
    np.testing.assert_allclose(out["k"][0][live], k2[live] / boxsize, rtol=1e-15)
    a.close(); b.close()


def test_output_before_the_first_cycle_ends_is_not_interpolated():
    """z_pk above z_in: written at cycle 0 with no look-ahead file."""
    N, numbins = 16, 8
    sim, (a, b), (sa, sb) = _boxes(N, 7, True)
    out = a.run_to_outputs(sim, sa, [150.0], numbins, 80.0)
    assert out["cycle"][0] == 0 and sa.cycle == 0
    b.cycle_fields(sim, sb)
    _, p, _, _, occ = _delta_power(b, sim, sb, numbins)
    rescalep = float(N) ** 6 * 2.0 * np.pi ** 2 * 0.31 * 0.31
    np.testing.assert_allclose(out["pk"][0][occ > 0], (p / rescalep)[occ > 0], rtol=1e-14)
    assert out["z"][0] == 1.0 / sa.a - 1.0
    a.close(); b.close()


def test_two_outputs_in_descending_order():
    N, numbins = 16, 8
    sim, (a, b), (sa, sb) = _boxes(N, 9, True)
    out = a.run_to_outputs(sim, sa, [98.0, 95.0], numbins, 80.0)
    assert out["cycle"][0] < out["cycle"][1] == sa.cycle
    assert np.allclose(out["z"], [98.0, 95.0], atol=1e-12)
    a.close(); b.close()
