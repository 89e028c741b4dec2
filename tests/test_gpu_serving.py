"""Serving loop on the GPU: each request's result is the same alone, in a
mix of concurrent requests, and on a reused box (bit for bit: every
request's kernels touch only its own box, in a fixed order)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "gevolution-1.2_amd"))

pytestmark = pytest.mark.gpu


def _requests():
    import serving as sv
    grids = [16, 32, 48, 16, 32, 48, 32]
    out = []
    for i, g in enumerate(grids):
        out.append(sv.Request(i, g, gr=(i % 2 == 0), Omega_m=0.26 + 0.015 * i, A_s=(1.9 + 0.07 * i) * 1e-9,
                              seed=100 + i))
    return out


def _same(a, b):
    assert a["cycles"] == b["cycles"] and a["a"] == b["a"]
    assert a["phi_sum"] == b["phi_sum"] and a["maxvel"] == b["maxvel"]
    np.testing.assert_array_equal(a["pk"], b["pk"])
    np.testing.assert_array_equal(a["occupation"], b["occupation"])


def test_results_alone_equal_results_in_the_mix_and_on_reused_boxes():
    import gevpm as gev
    import serving as sv
    tpl = gev.load_homogeneous_template(os.path.join(ROOT, "tests", "golden", "sc1_crystal.dat"))
    be = sv.GevBackend(template=tpl)
    reqs = _requests()
    alone = {}
    for r in reqs:
        with sv.Server(be, sv.default_budget(), slots=1, reuse=False) as s:
            (o,) = s.serve([r])
        assert o.status == "ok", o.error
        alone[r.rid] = o.result
        assert o.result["a"] >= 1.0 and o.result["cycles"] > 50 and o.result["z"] == 0.0
        assert np.all(np.isfinite(o.result["pk"])) and np.max(o.result["pk"]) > 0
    print("alone:", {k: (v["cycles"], v["phi_sum"]) for k, v in alone.items()}, flush=True)
    with sv.Server(be, sv.default_budget(), slots=4, reuse=True) as s:
        mix = s.serve(reqs)
        assert s.peak_running >= 2
    for o in mix:
        assert o.status == "ok", o.error
        _same(o.result, alone[o.rid])
    # one slot, so the repeated shapes are served on reused boxes
    with sv.Server(be, sv.default_budget(), slots=1, reuse=True) as s:
        seq = s.serve(reqs)
        assert s.created == 3
    assert sum(o.reused for o in seq) == len(reqs) - 3
    for o in seq:
        _same(o.result, alone[o.rid])


def test_served_run_matches_the_oracle():
    """A served request (IC generation, the main loop to z = 0 with the
    delta spectrum written as EXACT_OUTPUT_REDSHIFTS writes it) against the
    oracle's generator and output loop on the same inputs: same cycles,
    occupations exact, P(k) at z = 0 within 1e-6."""
    import gevpm as gev
    import serving as sv
    import bbks
    import oracle as O
    N = 32
    tpl = O.load_template(os.path.join(ROOT, "tests", "golden", "sc1_crystal.dat"))
    r = sv.Request(0, N, gr=True, Omega_m=0.3, A_s=2.1e-9, seed=5)
    with sv.Server(sv.GevBackend(template=tpl), sv.default_budget(), slots=1, reuse=False) as s:
        (o,) = s.serve([r])
    assert o.status == "ok", o.error
    L, h = r.box_mpc, r.h
    k, d, t = bbks.bbks_table(np.logspace(-5.0, np.log10(300.0), 400), 100.0, r.Omega_m, h)
    cosmo = O.make_cosmo(r.Omega_m - r.Omega_b, r.Omega_b, h)
    sim = O.Sim(N, 1, 0, 48.0, 0.04, 100.0, 1.5 * L * L / sv.C_LIGHT ** 2, cosmo)
    ob = O.OracleBox(N, N ** 3)
    ob.generate_ic(sim, O.IC(5, 1, 1, N // 4, 1, 100.0, 2.1e-9, 0.965, 0.05, L), tpl,
                   k * L, d, t * L / h, d, t * L / h)
    a0 = 1.0 / 101.0
    st = O.State(a0, 0.0, O.lib().orc_next_dtau(O.ctypes.byref(sim), a0), 0.0, 0.0, 0.0, 0.0, 0)
    want = ob.run_to_outputs(sim, st, [0.0], N // 2, L)
    ob.close()
    res = o.result
    assert res["cycles"] == st.cycle and res["a"] == st.a and res["z"] == want["z"][0] == 0.0
    np.testing.assert_array_equal(res["occupation"], want["occupation"][0])
    live = res["occupation"] > 0
    err = np.max(np.abs(res["pk"] - want["pk"][0])[live] / want["pk"][0][live])
    print("served 32^3 GR run: %d cycles, P(k) at z = 0 within %.2e of the oracle" % (st.cycle, err), flush=True)
    assert err < 1e-6


def test_continuous_batching_gives_every_request_its_plain_result():
    """BatchServer (requests of one shape share a batch and join it while
    others run) against each request run alone on a box of its own: the
    same cycles, phi sum and z = 0 spectrum, bit for bit."""
    import gevpm as gev
    import serving as sv
    tpl = gev.load_homogeneous_template(os.path.join(ROOT, "tests", "golden", "sc1_crystal.dat"))
    reqs = _requests()
    for k, r in enumerate(reqs):
        r.arrival = 0.01 * k  # later requests arrive while earlier ones run
    with sv.BatchServer(sv.BatchBackend(template=tpl), sv.default_budget(), slots=lambda key: 2) as s:
        served = s.serve(reqs)
        assert s.peak_running >= 3 and s.rounds > 100
    be = sv.GevBackend(template=tpl)
    for o, r in zip(served, reqs):
        assert o.status == "ok", o.error
        with gev.Box(r.ngrid, r.npart) as b:
            alone = be.run(b, r)
        _same(o.result, alone)
