"""The 512^3 GR configuration (box 2048 Mpc/h, BBKS LCDM ICs from seed 7,
parabolic B, FP64) through whole cycles on the GPU:

- the device IC generator's state is handed to the oracle and both run the
  first two cycles (cycle 0 without and cycle 1 with the phi update); phi,
  chi, B, positions and momenta agree within 1e-10 relative, per-cell
  counts bit for bit;
- a second box from the same seed runs the same cycles bit-identically
  (the run is reproducible run to run);
- the particle count is conserved, and T00hom stays within the kinetic and
  potential corrections of Omega_m;
- the FP32 form carried from the same state tracks the FP64 form.
Sizes: two FP64 boxes of 42 GB and one FP32 box of 30 GB on the device; the
oracle box holds about 32 GB of host memory."""
import os
import time

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
C_LIGHT = 2997.92458


def rel(a, b):
    s = float(np.max(np.abs(b)))
    d = float(np.max(np.abs(a - b)))
    return d / s if s > 0 else d


def test_c2_first_cycles_512(gev):
    import bbks
    N, L, h, seed = 512, 2048.0, 0.67, 7
    t_start = time.perf_counter()

    def say(msg):
        print("[c2 512^3 %.0f s] %s" % (time.perf_counter() - t_start, msg), flush=True)

    k, d, t = bbks.bbks_table(np.logspace(-5.0, np.log10(300.0), 400), 100.0, 0.31, h)
    tabs = (k * L, d, t * L / h, d, t * L / h)
    tpl = gev.load_homogeneous_template(os.path.join(ROOT, "tests", "golden", "sc1_crystal.dat"))
    fpg = 1.5 * L * L / C_LIGHT ** 2
    sim_g = gev.SimParams(1, 0, 48.0, 0.04, 100.0, fpg, gev.make_cosmology(0.26, 0.05, h))
    ic = gev.ICParams(seed, 1, 1, N // 4, 100.0, 2.1e-9, 0.965, 0.05, L)
    a0 = 1.0 / 101.0
    dt0 = gev.next_dtau(sim_g, N, a0)
    sim_o = O.Sim(N, 1, 0, 48.0, 0.04, 100.0, fpg, O.make_cosmo(0.26, 0.05, h))
    with gev.Box(N, N ** 3) as b, gev.Box(N, N ** 3) as b2:
        b.generateIC_basic(sim_g, ic, tpl, *tabs)
        b2.generateIC_basic(sim_g, ic, tpl, *tabs)
        say("ICs generated twice")
        ob = O.OracleBox(N, N ** 3)
        pos, q, ids = b.download_particles()
        order = np.argsort(ids)  # the oracle keeps its particles in this (ID) order
        ob.pos[:] = pos[order]
        ob.q[:] = q[order]
        ob.id[:] = ids[order]
        del pos, q, ids, order
        ob.box.pmass = 0.31 / N ** 3
        for f, name in ((gev.PHI, "phi"), (gev.CHI, "chi"), (gev.BI, "Bi"), (gev.BIFT, "BiFT"),
                        (gev.SCALARFT, "scalarFT")):
            getattr(ob, name)[:] = b.download_field(f)
        ob.file()
        say("state handed to the oracle")
        with gev.Box(N, N ** 3, form=gev.FORM_FP32) as b32:
            b32.copy_state_from(b)
            sg = gev.CycleState(a0, 0.0, dt0, 0.0, 0.0, 0.0, 0.0, 0, 0)
            sg2 = gev.CycleState(a0, 0.0, dt0, 0.0, 0.0, 0.0, 0.0, 0, 0)
            s32 = gev.CycleState(a0, 0.0, dt0, 0.0, 0.0, 0.0, 0.0, 0, 0)
            so = O.State(a0, 0.0, dt0, 0.0, 0.0, 0.0, 0.0, 0)
            for c in range(2):
                b.run_cycles(sim_g, sg, 1)
                b2.run_cycles(sim_g, sg2, 1)
                b32.run_cycles(sim_g, s32, 1)
                ob.cycle(sim_o, so)
                say("cycle %d done on the GPU (FP64 twice, FP32) and the oracle" % c)
                assert (sg.a, sg.tau, sg.dtau, sg.cycle) == (so.a, so.tau, so.dtau, so.cycle)
                assert sg2.a == sg.a and s32.a == sg.a
                if c == 0:
                    # T00hom = Omega_m up to the kinetic and potential terms (~1e-5 here)
                    assert abs(sg.T00hom / 0.31 - 1.0) < 1e-4, sg.T00hom
                    assert abs(sg.T00hom - so.T00hom) <= 1e-12 * so.T00hom
                errs = {}
                for f, name in ((gev.PHI, "phi"), (gev.CHI, "chi"), (gev.BI, "Bi")):
                    g = b.download_field(f)
                    errs[name] = rel(g, getattr(ob, name))
                    assert np.array_equal(g, b2.download_field(f)), (c, name)
                    g32 = b32.download_field(f)
                    errs[name + "_fp32"] = rel(g32, g)
                    del g, g32
                pos, q, ids = b.download_particles()
                pos2, q2, ids2 = b2.download_particles()
                assert np.array_equal(pos, pos2) and np.array_equal(q, q2) and np.array_equal(ids, ids2)
                del pos2, q2, ids2
                order = np.argsort(ids)
                dp = np.abs(pos[order] - ob.pos)
                errs["pos"] = float(np.minimum(dp, 1.0 - dp).max())
                errs["q"] = rel(q[order], ob.q)
                del pos, q, ids, order, dp
                counts, _ = b.download_filing()
                assert int(counts.sum()) == N ** 3
                errs["count_mismatch"] = int(np.count_nonzero(counts != ob.count))
                say("cycle %d: %s" % (c, {k: ("%.2e" % v if isinstance(v, float) else v) for k, v in errs.items()}))
                assert errs["count_mismatch"] == 0
                assert max(errs["phi"], errs["chi"], errs["Bi"], errs["q"]) < 1e-10 and errs["pos"] < 1e-12, errs
                assert errs["phi_fp32"] < 1e-4 and errs["chi_fp32"] < 1e-4, errs
        ob.close()
