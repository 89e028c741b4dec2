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
        assert o.result["a"] >= 1.0 and o.result["cycles"] > 50
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


def test_spectrum_of_a_served_run_equals_the_direct_calls():
    """The serving backend runs exactly the C-ABI sequence a driver would:
    IC generation, the cycles to z = 0, the delta power spectrum."""
    import gevpm as gev
    import serving as sv
    tpl = gev.load_homogeneous_template(os.path.join(ROOT, "tests", "golden", "sc1_crystal.dat"))
    r = sv.Request(0, 32, gr=True, Omega_m=0.3, A_s=2.1e-9, seed=5)
    with sv.Server(sv.GevBackend(template=tpl), sv.default_budget(), slots=1, reuse=False) as s:
        (o,) = s.serve([r])
    assert o.status == "ok", o.error
    import bbks
    sim, st = sv.GevBackend.setup(r)
    L, h = r.box_mpc, r.h
    k, d, t = bbks.bbks_table(np.logspace(-5.0, np.log10(300.0), 400), 100.0, r.Omega_m, h)
    with gev.Box(32, 32 ** 3) as b:
        ic = gev.ICParams(5, 1, 1, 8, 100.0, 2.1e-9, 0.965, 0.05, L)
        st.maxvel, _ = b.generateIC_basic(sim, ic, tpl, k * L, d, t * L / h, d, t * L / h)
        while 1.0 / st.a - 1.0 > 1e-12:
            b.run_cycles(sim, st, 1)
        b.projection_init(gev.SOURCE)
        b.projection_T00_project(st.a, True, 1.0)
        b.fft(gev.PLAN_SOURCE, gev.FFT_FORWARD)
        p = b.extractPowerSpectrum(gev.SCALARFT, 16)
        phi_sum = b.field_sum(gev.PHI)
    assert o.result["cycles"] == st.cycle and o.result["a"] == st.a
    np.testing.assert_array_equal(o.result["pk"], p[1])
    assert o.result["phi_sum"] == phi_sum
