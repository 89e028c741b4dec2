"""Batches (include/gev_pm.h gev_batch_*): boxes of one shape with their own
cosmologies, seeds and time steps, stepped by the same launches, through the
C ABI.  Every box's outputs and final state equal those of the same box run
alone through gev_run_to_outputs / gev_run_cycles bit for bit, with the
cycle's two stages replayed from HIP graphs or launched one by one; one box
is checked against the oracle (P(k) at z = 0 within 1e-6)."""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
C_LIGHT = 2997.92458


def cosmo_of(i):
    return 0.25 + 0.1 * ((i * 0.37) % 1.0), (1.8 + 0.6 * ((i * 0.61) % 1.0)) * 1e-9


def box_setup(gev, N, i, gr):
    """(sim, state, ICParams, tables, boxsize) of box i: Omega_m and A_s
    spread over [0.25, 0.35] and [1.8, 2.4]e-9, seed i, box 5 Mpc/h per cell."""
    import bbks
    Om, A_s = cosmo_of(i)
    L, h = 5.0 * N, 0.67
    k, d, t = bbks.bbks_table(np.logspace(-5.0, np.log10(300.0), 400), 100.0, Om, h)
    sim = gev.SimParams(gr, 0, 48.0, 0.04, 100.0, 1.5 * L * L / C_LIGHT ** 2, gev.make_cosmology(Om - 0.05, 0.05, h))
    a0 = 1.0 / 101.0
    st = gev.CycleState(a0, 0.0, gev.next_dtau(sim, N, a0), 0.0, 0.0, 0.0, 0.0, 0, 0)
    ic = gev.ICParams(i, 1, 1, N // 4, 100.0, A_s, 0.965, 0.05, L)
    return sim, st, ic, (k * L, d, t * L / h, d, t * L / h), L


def template(gev):
    return gev.load_homogeneous_template(os.path.join(ROOT, "tests", "golden", "sc1_crystal.dat"))


def generate(gev, box, N, i, gr, tpl, form):
    """ICs of box i into box (an FP64 box, or an FP32 one through a
    temporary FP64 box); returns (sim, state, boxsize)."""
    sim, st, ic, tabs, L = box_setup(gev, N, i, gr)
    if form == gev.FORM_FP64:
        box.generateIC_basic(sim, ic, tpl, *tabs)
    else:
        with gev.Box(N, N ** 3) as b64:
            b64.generateIC_basic(sim, ic, tpl, *tabs)
            box.copy_state_from(b64)
    return sim, st, L


def state_of(gev, box):
    pos, q, ids = box.download_particles()
    o = np.argsort(ids)
    return {"phi": box.download_field(gev.PHI), "chi": box.download_field(gev.CHI), "Bi": box.download_field(gev.BI),
            "pos": pos[o], "q": q[o]}


def equal_states(a, b):
    for k in a:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("gr,form,graph", [(1, 0, True), (0, 0, True), (1, 1, True), (1, 0, False)])
def test_batch_boxes_equal_boxes_run_alone(gev, gr, form, graph):
    N, nbox, numbins = 32, 4, 16
    z_pk = [10.0, 0.0]
    tpl = template(gev)
    with gev.Batch(N, N ** 3, nbox, gr=gr, form=form) as B:
        B.set_outputs(z_pk, numbins)
        for i in range(nbox):
            sim, st, L = generate(gev, B.box(i), N, i, gr, tpl, form)
            B.start_box(i, sim, st, L)
        B.run_to_outputs(graph=graph)
        for i in range(nbox):
            got = B.outputs(i)
            sg = B.state(i)
            gs = state_of(gev, B.box(i))
            with gev.Box(N, N ** 3, form=form) as b:
                sim, st, L = generate(gev, b, N, i, gr, tpl, form)
                want = b.run_to_outputs(sim, st, z_pk, numbins, L)
                ws = state_of(gev, b)
            assert (sg.a, sg.tau, sg.dtau, sg.dtau_old, sg.cycle, sg.stage) == \
                (st.a, st.tau, st.dtau, st.dtau_old, st.cycle, st.stage), i
            assert sg.maxvel == st.maxvel, i
            for k in ("k", "pk", "sigma_k", "sigma_pk", "occupation", "z", "cycle", "interpolated"):
                assert np.array_equal(got[k], want[k]), (i, k)
            equal_states(gs, ws)
        print("gr=%d form=%d graph=%d: %d boxes equal alone; cycles %s" % (gr, form, graph, nbox,
              [int(B.state(i).cycle) for i in range(nbox)]), flush=True)


def test_batch_run_cycles_from_different_starts_and_restarted_slots(gev):
    """Boxes join at different cycles (a slot is reused after a stop), plain
    cycles without outputs: each equals its box run alone."""
    N = 16
    tpl = template(gev)
    with gev.Batch(N, N ** 3, 3, gr=1) as B:
        for i in range(3):
            sim, st, L = generate(gev, B.box(i), N, i, 1, tpl, 0)
            B.start_box(i, sim, st, L)
        B.run_cycles(5)
        B.stop_box(1)
        sim1, st1, L1 = generate(gev, B.box(1), N, 7, 1, tpl, 0)  # a new run in slot 1
        B.start_box(1, sim1, st1, L1)
        B.run_cycles(4, graph=False)
        B.run_cycles(3, graph=True)
        for i, seed, n in ((0, 0, 12), (1, 7, 7), (2, 2, 12)):
            got = state_of(gev, B.box(i))
            with gev.Box(N, N ** 3) as b:
                sim, st, L = generate(gev, b, N, seed, 1, tpl, 0)
                b.run_cycles(sim, st, n)
                equal_states(got, state_of(gev, b))
            assert B.state(i).a == st.a and B.state(i).cycle == n


def test_batch_box_against_the_oracle(gev):
    N, numbins = 32, 16
    tpl = template(gev)
    tplo = O.load_template(os.path.join(ROOT, "tests", "golden", "sc1_crystal.dat"))
    with gev.Batch(N, N ** 3, 2, gr=1) as B:
        B.set_outputs([0.0], numbins)
        for i in range(2):
            sim, st, L = generate(gev, B.box(i), N, i + 3, 1, tpl, 0)
            B.start_box(i, sim, st, L)
        B.run_to_outputs()
        got = B.outputs(1)
        cyc = B.state(1).cycle
    sim, st, ic, tabs, L = box_setup(gev, N, 4, 1)
    Om, A_s = cosmo_of(4)
    so = O.Sim(N, 1, 0, 48.0, 0.04, 100.0, sim.fourpiG, O.make_cosmo(Om - 0.05, 0.05, 0.67))
    ob = O.OracleBox(N, N ** 3)
    ob.generate_ic(so, O.IC(4, 1, 1, N // 4, 1, 100.0, A_s, 0.965, 0.05, L), tplo, *tabs)
    s0 = O.State(st.a, 0.0, st.dtau, 0.0, 0.0, 0.0, 0.0, 0)
    want = ob.run_to_outputs(so, s0, [0.0], numbins, L)
    ob.close()
    assert cyc == s0.cycle and np.array_equal(got["occupation"], want["occupation"])
    live = want["occupation"][0] > 0
    err = np.max(np.abs(got["pk"][0] - want["pk"][0])[live] / want["pk"][0][live])
    print("batch box vs oracle: %d cycles, P(k) at z = 0 within %.2e" % (cyc, err), flush=True)
    assert err < 1e-6


def test_ensemble_of_two_grid_sizes_equals_boxes_alone(gev):
    """The ensemble driver (one batch per grid size, stepped round-robin on
    their own streams) gives every box the result of the box run alone."""
    import ensemble as en
    tpl = template(gev)
    specs = en.draw(6, grids=(16, 32), seed=11)
    with en.Ensemble(specs, tpl, form=gev.FORM_FP32) as E:
        E.generate()
        E.run()
        res = [E.result(s.index) for s in specs]
    for s, r in zip(specs, res):
        sim, st, ic, tabs = en.setup(s)
        with gev.Box(s.ngrid, s.ngrid ** 3) as b64, gev.Box(s.ngrid, s.ngrid ** 3, form=gev.FORM_FP32) as b:
            b64.generateIC_basic(sim, ic, tpl, *tabs)
            b.copy_state_from(b64)
            out = b.run_to_outputs(sim, st, [0.0], s.ngrid // 2, s.boxsize)
        assert r["cycles"] == st.cycle and r["a"] == st.a and r["z"] == 0.0
        assert np.array_equal(r["pk"], out["pk"][0]) and np.array_equal(r["occupation"], out["occupation"][0])
