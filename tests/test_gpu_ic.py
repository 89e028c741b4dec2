"""GPU parity of the basic IC generator (ic_basic.hpp) against the oracle.

Bit-exact: PRNG outputs (golden vectors of the reference's own engine), the
CIC kernel, the particle lattice, the IC displacement and momenta for a given
field.  Tolerances: the Gaussian realisation uses the device cos/sin, whose
last bit may differ from the C library's (1e-13 relative per mode); the whole
generator, which adds FFT rounding, within 1e-10 relative (the tolerance of
the per-step fields)."""
import json
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
C_LIGHT = 2997.92458  # metadata.hpp:98, units of 100 km/s


def rel(a, b):
    d = np.max(np.abs(np.asarray(a) - np.asarray(b)))
    s = np.max(np.abs(np.asarray(b)))
    return d / s if s > 0 else d


def test_prng_fill_matches_reference_golden_vectors(gev):
    g = json.load(open(os.path.join(GOLD, "sitmo_prng.json")))
    with gev.Box(8, 8) as b:
        for c in g["cases"]:
            pos, got = 0, []
            for st in c["steps"]:
                n = int(st[1:])
                if st[0] == "d":
                    got += b.prng_fill(c["seed"], pos, n).tolist()
                pos += n
            assert got == c["values"], (c["seed"], c["steps"])
        # a long run against the oracle, across many blocks
        v = b.prng_fill(42, (3 * 65536 + 5) * 65536 * 65536 + 3, 4096)
        assert all(int(v[i]) == O.prng_at(42, (3 * 65536 + 5) * 65536 * 65536 + 3 + i) for i in range(0, 4096, 37))


def unit_spline():
    return np.array([0.0, 1e6]), np.array([1.0, 1.0])


@pytest.mark.parametrize("N,seed,ksphere", [(16, 42, 1), (16, 7, 0), (12, 3, 1), (32, 5, 1)])
def test_realisation_unit_spectrum(gev, N, seed, ksphere):
    """Realisation with K = 1, P = 1: the draws and their placement."""
    x, y = unit_spline()
    with gev.Box(N, 8) as b:
        b.generateDisplacementField(0.0, x, y, seed, ksphere, deconvolve_f=0, ignorekernel=1)
        got = b.download_field(gev.SCALARFT)
    want = O.generate_displacement_field(N, np.zeros_like(got), 0.0, x, y, O.spline_init(x, y), seed, ksphere,
                                         deconvolve_f=0, ignorekernel=1, cr_log=1)
    assert np.array_equal(got == 0, want == 0)
    nz = want != 0
    assert np.max(np.abs(got[nz] - want[nz]) / np.abs(want[nz])) < 1e-13


def test_realisation_zero_redraw_row(gev):
    g = json.load(open(os.path.join(GOLD, "prng_rejection.json")))
    N, seed = g["N"], g["seed"]
    x, y = unit_spline()
    with gev.Box(N, 8) as b:
        b.generateDisplacementField(0.0, x, y, seed, g["ksphere"], deconvolve_f=0, ignorekernel=1)
        got = b.download_field(gev.SCALARFT)
    want = O.generate_displacement_field(N, np.zeros_like(got), 0.0, x, y, O.spline_init(x, y), seed,
                                         g["ksphere"], deconvolve_f=0, ignorekernel=1, cr_log=1)
    nz = want != 0
    assert np.array_equal(got == 0, want == 0)
    assert np.max(np.abs(got[nz] - want[nz]) / np.abs(want[nz])) < 1e-13


@pytest.mark.parametrize("name,ratio", [("sc1", 4), ("bcc", 4), ("fcc", 2), (None, 1)])
def test_cic_kernel_and_deconvolved_realisation(gev, name, ratio):
    N = 32
    tpl = O.load_template(os.path.join(GOLD, name + "_crystal.dat")) if name else None
    nt = N // ratio
    x = np.linspace(1.0, 2 * np.pi * N, 300)
    y = 1e-5 / (1.0 + 0.01 * x ** 2)
    with gev.Box(N, 8) as b:
        b.generateCICKernel(gev.SOURCE, tpl, nt)
        ker = b.download_field(gev.SOURCE)
        want = O.generate_cic_kernel(N, tpl, nt) if name else O.generate_cic_kernel(N)
        assert np.array_equal(ker, want)
        b.fft(gev.PLAN_SOURCE, gev.FFT_FORWARD)
        kft = b.download_field(gev.SCALARFT)
        b.generateDisplacementField(0.3, x, y, 9, 1)
        got = b.download_field(gev.SCALARFT)
    F = O.generate_displacement_field(N, kft, 0.3, x, y, O.spline_init(x, y), 9, 1, cr_log=1)
    nz = F != 0
    assert np.array_equal(got == 0, F == 0)
    assert np.max(np.abs(got[nz] - F[nz]) / np.abs(F[nz])) < 1e-13


def test_particle_lattice_displacement_and_momenta(gev):
    N = 32
    tpl = O.load_template(os.path.join(GOLD, "sc1_crystal.dat"))
    nt = N // 4
    rng = np.random.default_rng(4)
    z, yy, xx = np.meshgrid(*(np.arange(N) / N,) * 3, indexing="ij")
    f = 1e-3 * (np.sin(2 * np.pi * xx) + np.cos(4 * np.pi * yy) * np.sin(2 * np.pi * z)) + \
        1e-5 * rng.standard_normal((N, N, N))
    pos_o, ids_o = O.initialize_particle_positions(tpl, nt)
    with gev.Box(N, N ** 3) as b:
        b.initializeParticlePositions(tpl, nt, 0.31 / N ** 3)
        pos, q, ids = b.download_particles()
        assert np.array_equal(ids, ids_o) and np.array_equal(pos, pos_o) and not q.any()
        counts, _ = b.download_filing()
        assert counts.sum() == N ** 3 and counts.max() == 1
        b.upload_field(gev.CHI, f)
        md = b.displace_pcls_ic(0.7)
        mo = O.displace_pcls_ic(N, pos_o, 0.7, f)
        pos, _, ids = b.download_particles()
        assert np.array_equal(pos, pos_o) and md == mo
        counts, cells = b.download_filing()
        ocell, oc, _, _ = O.file_particles(N, pos_o, ids_o)
        assert np.array_equal(counts, oc) and np.array_equal(cells, ocell)
        mv = b.initialize_q_ic(1.3)
        qo, mvo = O.initialize_q_ic(N, pos_o, 1.3, f)
        _, q, _ = b.download_particles()
        assert np.array_equal(q, qo) and mv == mvo


def ic_inputs(N, boxsize, gr, seed=42):
    cosmo_o = O.make_cosmo()
    fpg = 1.5 * boxsize ** 2 / C_LIGHT ** 2
    tk = os.path.join(GOLD, "class_tk.dat")
    k, dt, tt = O.load_transfer_functions(tk, "tot", boxsize, cosmo_o.h)
    _, dc, tc = O.load_transfer_functions(tk, "cdm", boxsize, cosmo_o.h)
    tpl = O.load_template(os.path.join(GOLD, "sc1_crystal.dat"))
    sim_o = O.Sim(N, gr, 0, 48.0, 0.04, 30.0, fpg, cosmo_o)
    ic_o = O.IC(seed, 1, 1, N // 4, 1, 100.0, 2.1e-9, 0.965, 0.05, boxsize)
    return cosmo_o, fpg, (k, dt, tt, dc, tc), tpl, sim_o, ic_o


@pytest.mark.parametrize("N,gr", [(16, 1), (32, 1), (32, 0), (48, 1)])
def test_generate_ic_basic_matches_oracle(gev, N, gr):
    boxsize = 5.0 * N
    cosmo_o, fpg, tabs, tpl, sim_o, ic_o = ic_inputs(N, boxsize, gr)
    ob = O.OracleBox(N, N ** 3)
    mvo, mdo = ob.generate_ic(sim_o, ic_o, tpl, *tabs)
    sim_g = gev.SimParams(gr, 0, 48.0, 0.04, 30.0, fpg, gev.make_cosmology())
    ic_g = gev.ICParams(ic_o.seed, 1, 1, N // 4, 100.0, 2.1e-9, 0.965, 0.05, boxsize)
    with gev.Box(N, N ** 3) as b:
        mv, md = b.generateIC_basic(sim_g, ic_g, tpl, *tabs)
        for name, f in (("phi", gev.PHI), ("chi", gev.CHI), ("Bi", gev.BI), ("BiFT", gev.BIFT)):
            g = b.download_field(f)
            r = getattr(ob, name)
            assert rel(g, r) < 1e-10, (name, rel(g, r))
        pos, q, ids = b.download_particles()
        assert np.array_equal(ids, ob.id)
        assert rel(pos, ob.pos) < 1e-12 and rel(q, ob.q) < 1e-10
        counts, cells = b.download_filing()
        assert np.array_equal(counts, ob.count) and np.array_equal(cells, ob.cell)
        assert abs(md - mdo) <= 1e-10 * mdo and abs(mv - mvo) <= 1e-10 * mvo
    ob.close()


def test_generate_ic_then_cycles(gev):
    """The generator's output state runs through gev_run_cycles like the
    oracle's through orc_cycle."""
    N, boxsize = 32, 160.0
    cosmo_o, fpg, tabs, tpl, sim_o, ic_o = ic_inputs(N, boxsize, 1)
    ob = O.OracleBox(N, N ** 3)
    ob.generate_ic(sim_o, ic_o, tpl, *tabs)
    ob.file()
    sim_g = gev.SimParams(1, 0, 48.0, 0.04, 30.0, fpg, gev.make_cosmology())
    ic_g = gev.ICParams(ic_o.seed, 1, 1, N // 4, 100.0, 2.1e-9, 0.965, 0.05, boxsize)
    a0 = 1.0 / 101.0
    dt0 = O.lib().orc_next_dtau(O.ctypes.byref(sim_o), a0)
    so = O.State(a0, O.particle_horizon(a0, fpg, cosmo_o), dt0, 0.0, 0.0, 0.0, 0.0, 0)
    with gev.Box(N, N ** 3) as b:
        b.generateIC_basic(sim_g, ic_g, tpl, *tabs)
        sg = gev.CycleState(a0, so.tau, dt0, 0.0, 0.0, 0.0, 0.0, 0, 0)
        for c in range(3):
            ob.cycle(sim_o, so)
            b.run_cycles(sim_g, sg, 1)
            assert so.a == sg.a
            for name, f in (("phi", gev.PHI), ("chi", gev.CHI), ("Bi", gev.BI)):
                assert rel(b.download_field(f), getattr(ob, name)) < 1e-9, (c, name)
            pos, q, ids = b.download_particles()
            order = np.argsort(ids)
            assert rel(pos[order], ob.pos) < 1e-12 and rel(q[order], ob.q) < 1e-9
    ob.close()
