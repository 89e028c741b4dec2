"""Compact boxes (GEV_BOX_COMPACT: one particle set, drifts written in place)
through the C ABI.  Every deposit sums a cell's particles in ascending ID
whatever their storage order, so a compact box's fields, positions and
momenta equal the plain box's bit for bit; at 1024^3 the FP64 form fits one
MI355X only compact (277 GB against 337 GB)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
C_LIGHT = 2997.92458


def ic(gev, box, N, L, seed):
    import bbks
    h = 0.67
    k, d, t = bbks.bbks_table(np.logspace(-5.0, np.log10(300.0), 400), 100.0, 0.31, h)
    sim = gev.SimParams(1, 0, 48.0, 0.04, 100.0, 1.5 * L * L / C_LIGHT ** 2, gev.make_cosmology(0.26, 0.05, h))
    a0 = 1.0 / 101.0
    st = gev.CycleState(a0, 0.0, gev.next_dtau(sim, N, a0), 0.0, 0.0, 0.0, 0.0, 0, 0)
    tpl = gev.load_homogeneous_template(os.path.join(ROOT, "tests", "golden", "sc1_crystal.dat"))
    box.generateIC_basic(sim, gev.ICParams(seed, 1, 1, N // 4, 100.0, 2.1e-9, 0.965, 0.05, L), tpl,
                         k * L, d, t * L / h, d, t * L / h)
    return sim, st


def by_id(box):
    pos, q, ids = box.download_particles()
    o = np.argsort(ids)
    return pos[o], q[o]


@pytest.mark.parametrize("form", [0, 1])
def test_compact_box_equals_plain_box(gev, form):
    N, L = 32, 160.0
    with gev.Box(N, N ** 3) as a, gev.Box(N, N ** 3, compact=True) as c64, \
            gev.Box(N, N ** 3, form=form) as b, gev.Box(N, N ** 3, form=form, compact=True) as c:
        sim, st = ic(gev, a, N, L, 3)
        sim2, st2 = ic(gev, c64, N, L, 3)
        b.copy_state_from(a)
        c.copy_state_from(c64)
        b.run_to_outputs(sim, st, [20.0], 16, L)
        pk = c.run_to_outputs(sim2, st2, [20.0], 16, L)["pk"]
        b.run_cycles(sim, st, 5)
        c.run_cycles(sim2, st2, 5)
        assert (st.a, st.cycle, st.maxvel) == (st2.a, st2.cycle, st2.maxvel)
        for f in (gev.PHI, gev.CHI, gev.BI, gev.BIFT):
            assert np.array_equal(b.download_field(f), c.download_field(f)), f
        pb, qb = by_id(b)
        pc, qc = by_id(c)
        assert np.array_equal(pb, pc) and np.array_equal(qb, qc)
        assert np.array_equal(b.download_filing()[0], c.download_filing()[0])
        assert np.all(np.isfinite(pk))


def test_compact_fp64_box_at_1024():
    """The 1024^3 FP64 box (box 4096 Mpc/h, seed 11) in its compact form:
    generated ICs and two cycles; particle count conserved, T00hom within
    the kinetic and potential corrections of Omega_m, fields finite."""
    import gevpm as gev
    N, L = 1024, 4096.0
    assert gev.footprint_bytes(N, N ** 3, gev.FORM_FP64, compact=True) < 280e9
    with gev.Box(N, N ** 3, compact=True) as b:
        sim, st = ic(gev, b, N, L, 11)
        b.run_cycles(sim, st, 2)
        assert abs(st.T00hom / 0.31 - 1.0) < 1e-4 and st.cycle == 2
        counts, _ = b.download_filing()
        assert int(counts.sum(dtype=np.int64)) == N ** 3
        del counts
        phi = b.download_field(gev.PHI)
        assert np.all(np.isfinite(phi)) and np.max(np.abs(phi)) > 0
        print("1024^3 compact FP64: T00hom %.8f after 2 cycles, max |phi| %.3e" % (st.T00hom, np.max(np.abs(phi))),
              flush=True)
