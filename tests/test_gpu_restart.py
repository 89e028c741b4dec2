"""Hibernation and restart (hibernation.hpp:512-612, ic_read.hpp): a run that
hibernates after k cycles and restarts in a fresh context continues like the
uninterrupted run (manual sec. 3.4: "exactly as if no hibernation had
occurred ... up to finite machine precision").  With the parabolic vector
method the restart rebuilds BiFT from B by a forward FFT, so B carries FFT
rounding (1e-13 relative); with the elliptic method the continuation is
bit-identical."""
import numpy as np
import pytest

from states import gr_state, sim_params

pytestmark = pytest.mark.gpu


def rel(a, b):
    s = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / s if s > 0 else np.max(np.abs(a - b))


def new_box(gev, st, N):
    b = gev.Box(N, len(st["ids"]))
    b.upload_particles(st["pos"], st["q"], st["ids"], st["pmass"])
    for f, k in ((gev.PHI, "phi"), (gev.CHI, "chi"), (gev.BI, "Bi"), (gev.BIFT, "BiFT"),
                 (gev.SCALARFT, "scalarFT")):
        b.upload_field(f, st[k])
    return b


@pytest.mark.parametrize("gr,elliptic,N", [(True, False, 32), (True, True, 32), (False, False, 32), (True, False, 48)])
def test_restart_continues_the_run(gev, tmp_path, gr, elliptic, N):
    st = gr_state(N, 17)
    sp = sim_params(N, 5.0 * N, gr=gr, vector_elliptic=elliptic)
    sim = gev.SimParams(sp["gr_flag"], sp["vector_flag"], sp["Cf"], sp["steplimit"], sp["z_relax"], sp["fourpiG"],
                        gev.make_cosmology())
    a0 = 1.0 / (1.0 + sp["z_in"])
    path = str(tmp_path / "hib.bin")
    with new_box(gev, st, N) as A:
        sa = gev.CycleState(a0, 0.0, gev.next_dtau(sim, N, a0), 0.0, 0.0, 0.0, 0.0, 0, 0)
        A.run_cycles(sim, sa, 3)
        A.hibernate(sim, sa, path)
        A.run_cycles(sim, sa, 2)
        pa, qa, ia = A.download_particles()
        fa = {f: A.download_field(f) for f in (gev.PHI, gev.CHI, gev.BI)}
    with gev.Box(N, len(st["ids"])) as B:
        sb = B.restart(sim, path)
        assert sb.cycle == 3
        B.run_cycles(sim, sb, 2)
        pb, qb, ib = B.download_particles()
        fb = {f: B.download_field(f) for f in (gev.PHI, gev.CHI, gev.BI)}
    assert sb.cycle == sa.cycle and sb.a == sa.a and sb.tau == sa.tau
    assert abs(sb.dtau - sa.dtau) <= 4e-16 * sa.dtau  # Cf/N on restart, Cf*dx in the loop
    oa, ob = np.argsort(ia), np.argsort(ib)
    assert np.array_equal(ia[oa], ib[ob])
    if elliptic:
        assert np.array_equal(pa[oa], pb[ob]) and np.array_equal(qa[oa], qb[ob])
        for f in fa:
            assert np.array_equal(fa[f], fb[f]), f
    else:
        assert rel(pb[ob], pa[oa]) < 1e-13 and rel(qb[ob], qa[oa]) < 1e-10
        for f in fa:
            if gr or f == gev.PHI:
                assert rel(fb[f], fa[f]) < 1e-10, (f, rel(fb[f], fa[f]))


def test_restart_rejects_a_mismatched_context(gev, tmp_path):
    N = 16
    st = gr_state(N, 2)
    sp = sim_params(N, 80.0)
    sim = gev.SimParams(1, 0, sp["Cf"], sp["steplimit"], sp["z_relax"], sp["fourpiG"], gev.make_cosmology())
    path = str(tmp_path / "h.bin")
    with new_box(gev, st, N) as A:
        s = gev.CycleState(0.01, 0.0, 1e-3, 0.0, 0.0, 0.0, 0.0, 0, 0)
        A.run_cycles(sim, s, 1)
        A.hibernate(sim, s, path)
    with gev.Box(24, N ** 3) as B:
        with pytest.raises(gev.GevError):
            B.restart(sim, path)
    with gev.Box(N, N ** 3) as B:
        with pytest.raises(gev.GevError):
            B.restart(gev.SimParams(1, 1, sp["Cf"], sp["steplimit"], sp["z_relax"], sp["fourpiG"],
                                    gev.make_cosmology()), path)  # other vector method


def test_background_records_match_the_oracle_cycles(gev, orc, tmp_path):
    """gev_run_cycles_recorded gives main.cpp:520-537's line per cycle: the
    cycle's starting tau and a, Hconf(a)/Hconf(1), phi(k=0) after the
    Poisson solve and T00hom, as the oracle's cycle records them."""
    N = 32
    st = gr_state(N, 9)
    sp = sim_params(N, 160.0)
    cosmo_o = orc.make_cosmo()
    sim_o = orc.Sim(N, 1, 0, sp["Cf"], sp["steplimit"], sp["z_relax"], sp["fourpiG"], cosmo_o)
    a0 = 1.0 / (1.0 + sp["z_in"])
    dt0 = orc.lib().orc_next_dtau(orc.ctypes.byref(sim_o), a0)
    so = orc.State(a0, 0.0, dt0, 0.0, 0.0, 0.0, 0.0, 0)
    ob = orc.OracleBox(N, len(st["ids"]))
    ob.pos[:] = st["pos"]; ob.q[:] = st["q"]; ob.id[:] = st["ids"]; ob.box.pmass = st["pmass"]
    ob.phi[:] = st["phi"]; ob.chi[:] = st["chi"]; ob.Bi[:] = st["Bi"]
    ob.BiFT[:] = st["BiFT"]; ob.scalarFT[:] = st["scalarFT"]
    ob.file()
    want = []
    for c in range(4):
        t0, a_0 = so.tau, so.a
        ob.cycle(sim_o, so)
        want.append((c, t0, a_0, so.phi_k0, so.T00hom))
    ob.close()
    sim = gev.SimParams(1, 0, sp["Cf"], sp["steplimit"], sp["z_relax"], sp["fourpiG"], gev.make_cosmology())
    with new_box(gev, st, N) as b:
        sg = gev.CycleState(a0, 0.0, dt0, 0.0, 0.0, 0.0, 0.0, 0, 0)
        recs = b.run_cycles_recorded(sim, sg, 4)
    H1 = gev.hconf(1.0, sp["fourpiG"], gev.make_cosmology())
    for r, (c, t0, a_0, phik0, t00) in zip(recs, want):
        assert r.cycle == c and r.tau == t0 and r.a == a_0
        assert r.hconf_ratio == gev.hconf(a_0, sp["fourpiG"], gev.make_cosmology()) / H1
        assert abs(r.phi_k0 - phik0) <= 1e-10 * abs(phik0) + 1e-300
        assert abs(r.T00hom - t00) <= 1e-12 * abs(t00)
    path = str(tmp_path / "run_background.dat")
    gev.write_background(path, recs)
    text = open(path).read().splitlines()
    assert text[0] == "# background statistics"
    rows = np.loadtxt(path)
    assert rows.shape == (4, 6) and np.array_equal(rows[:, 0], np.arange(4))
    assert np.allclose(rows[:, 5], [w[4] for w in want], rtol=1e-6)
