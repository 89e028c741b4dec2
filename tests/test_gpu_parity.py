"""GPU parity: every hot-path entry point of libgevpm.so against the CPU oracle.

Tolerances (from the task's parity bar):
  * bit-exact (np.array_equal) for integer work (cell indices, per-cell counts)
    and for kernels whose inputs are identical on both sides and whose
    arithmetic is the reference's expression sequence (stencils, k-space
    solves, deposits, kicks, drifts) -- both sides are built without FMA
    contraction;
  * FFTs: max |gpu - ref| <= 1e-13 * max |ref| (different algorithms);
  * whole cycles: phi, chi, Bi and positions within 1e-10 relative (max norm),
    cell counts bit-exact.
"""
import numpy as np
import pytest

from states import gr_state, lattice_particles, smooth_field, sim_params

pytestmark = pytest.mark.gpu

AX = (-3, -2, -1)


def rel(a, b):
    d = np.max(np.abs(np.asarray(a) - np.asarray(b)))
    s = np.max(np.abs(b))
    return d / s if s > 0 else d


def new_box(gev, st, N):
    b = gev.Box(N, len(st["ids"]))
    b.upload_particles(st["pos"], st["q"], st["ids"], st["pmass"])
    for f, key in ((gev.PHI, "phi"), (gev.CHI, "chi"), (gev.BI, "Bi"), (gev.BIFT, "BiFT"),
                   (gev.SCALARFT, "scalarFT")):
        b.upload_field(f, st[key])
    return b


# ----------------------------------------------------------------- FFT

@pytest.mark.parametrize("N", [8, 12, 16, 24, 32, 48, 64, 96, 128])
def test_fft_r2c_c2r_vs_numpy(gev, N):
    rng = np.random.default_rng(N)
    f = rng.standard_normal((N, N, N))
    with gev.Box(N, 0) as b:
        b.upload_field(gev.SOURCE, f)
        b.fft(gev.PLAN_SOURCE, gev.FFT_FORWARD)
        k = b.download_field(gev.SCALARFT)
        assert rel(k, np.fft.rfftn(f)) < 1e-13
        kk = rng.standard_normal(k.shape) + 1j * rng.standard_normal(k.shape)
        b.upload_field(gev.SCALARFT, kk)
        b.fft(gev.PLAN_PHI, gev.FFT_BACKWARD)
        back = b.download_field(gev.PHI)
        assert rel(back, np.fft.irfftn(kk, s=(N, N, N), axes=AX) * N ** 3) < 1e-13
        # the backward transform leaves its k input untouched (main.cpp:586/593)
        assert np.array_equal(b.download_field(gev.SCALARFT), kk)


def test_fft_multicomponent_and_roundtrip(gev):
    N = 32
    rng = np.random.default_rng(5)
    S = rng.standard_normal((6, N, N, N))
    with gev.Box(N, 0) as b:
        b.upload_field(gev.SIJ, S)
        b.fft(gev.PLAN_SIJ, gev.FFT_FORWARD)
        k = b.download_field(gev.SIJFT)
        assert rel(k, np.fft.rfftn(S, axes=AX)) < 1e-13
        B = rng.standard_normal((3, N, N, N))
        b.upload_field(gev.BI, B)
        b.fft(gev.PLAN_BI, gev.FFT_FORWARD)
        b.fft(gev.PLAN_BI, gev.FFT_BACKWARD)
        assert rel(b.download_field(gev.BI), B * N ** 3) < 1e-13


# ----------------------------------------------------------------- k space

def rand_k(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("N,modif", [(16, 0.0), (24, 0.37), (32, 2.5)])
def test_poisson_bit_exact(gev, orc, N, modif):
    rng = np.random.default_rng(1)
    s = rand_k(rng, (N, N, N // 2 + 1))
    with gev.Box(N, 0) as b:
        b.upload_field(gev.SCALARFT, s)
        b.solveModifiedPoissonFT(0.731, modif)
        assert np.array_equal(b.download_field(gev.SCALARFT), orc.solve_modified_poisson_ft(s, 0.731, modif))


@pytest.mark.parametrize("N", [16, 24])
def test_sij_solves_bit_exact(gev, orc, N):
    rng = np.random.default_rng(2)
    S = rand_k(rng, (6, N, N, N // 2 + 1))
    B0 = rand_k(rng, (3, N, N, N // 2 + 1))
    chi0 = rand_k(rng, (N, N, N // 2 + 1))
    with gev.Box(N, 0) as b:
        b.upload_field(gev.SIJFT, S)
        b.upload_field(gev.BIFT, B0)
        b.upload_field(gev.SCALARFT, chi0)
        b.projectFTscalar(1)
        assert np.array_equal(b.download_field(gev.SCALARFT), orc.project_ft_scalar(S, chi0, add=1))
        b.projectFTscalar(0)
        chi_ref = orc.project_ft_scalar(S)
        assert np.array_equal(b.download_field(gev.SCALARFT), chi_ref)
        b.evolveFTvector(0.0123)
        B_ref = orc.evolve_ft_vector(S, B0, 0.0123)
        assert np.array_equal(b.download_field(gev.BIFT), B_ref)
        # fused form equals the two separate calls
        b.upload_field(gev.BIFT, B0)
        b.upload_field(gev.SCALARFT, chi0)
        b.projectFTscalar_evolveFTvector(0.0123)
        assert np.array_equal(b.download_field(gev.SCALARFT), chi_ref)
        assert np.array_equal(b.download_field(gev.BIFT), B_ref)
        b.projectFTvector(0.5, 0.25)
        assert np.array_equal(b.download_field(gev.BIFT), orc.project_ft_vector(B_ref, 0.5, 0.25))


# ----------------------------------------------------------------- stencils

@pytest.mark.parametrize("N", [8, 24, 32])
def test_prepare_ft_source_bit_exact(gev, orc, N):
    rng = np.random.default_rng(3)
    phi = smooth_field(N, rng, 1e-4)
    chi = smooth_field(N, rng, 1e-5)
    src = 0.31 + smooth_field(N, rng, 0.5)
    T = smooth_field(N, rng, 1e-3, ncomp=6)
    with gev.Box(N, 0) as b:
        b.upload_field(gev.PHI, phi)
        b.upload_field(gev.CHI, chi)
        b.upload_field(gev.SOURCE, src)
        b.upload_field(gev.SIJ, T)
        b.prepareFTsource_scalar(0.31, 0.3, 0.02, 0.7)
        ref = orc.prepare_ft_source_scalar(phi, chi, src, 0.31, 0.3, 0.02, 0.7)
        assert np.array_equal(b.download_field(gev.SOURCE), ref)
        b.prepareFTsource_tensor(0.05)
        assert np.array_equal(b.download_field(gev.SIJ), orc.prepare_ft_source_tensor(phi, T, 0.05))


# ----------------------------------------------------------------- filing / deposit

def check_filing(gev, orc, b, N, pos, ids):
    cell_r, count_r, _, _ = orc.file_particles(N, pos, ids)
    counts, cells = b.download_filing()
    assert np.array_equal(counts, count_r)
    dpos, _, dids = b.download_particles()
    order = np.argsort(dids)
    assert np.array_equal(cells[order], cell_r[np.argsort(ids)])


@pytest.mark.parametrize("N,npc", [(16, 1), (24, 1), (16, 2)])
def test_deposits_bit_exact(gev, orc, N, npc):
    st = gr_state(N, 11, npc_side=npc)
    a = 0.0123
    with new_box(gev, st, N) as b:
        check_filing(gev, orc, b, N, st["pos"], st["ids"])
        args = (N, st["pos"], st["q"], st["ids"], st["pmass"])
        # T00 with and without phi
        b.projection_init(gev.SOURCE)
        b.projection_T00_project(a, True, 1.0)
        assert np.array_equal(b.download_field(gev.SOURCE), orc.projection_T00_project(*args, a=a, phi=st["phi"]))
        b.projection_init(gev.SOURCE)
        b.projection_T00_project(a, False, 0.7)
        assert np.array_equal(b.download_field(gev.SOURCE), orc.projection_T00_project(*args, a=a, coeff=0.7))
        # Tij
        b.projection_init(gev.SIJ)
        b.projection_Tij_project(a, True, 1.0)
        tij_ref = orc.projection_Tij_project(*args, a=a, phi=st["phi"])
        assert np.array_equal(b.download_field(gev.SIJ), tij_ref)
        # fused T00 + Tij
        b.projection_init(gev.SOURCE)
        b.projection_init(gev.SIJ)
        b.projection_T00_Tij_project(a, 1.0)
        assert np.array_equal(b.download_field(gev.SOURCE), orc.projection_T00_project(*args, a=a, phi=st["phi"]))
        assert np.array_equal(b.download_field(gev.SIJ), tij_ref)
        # T0i (into Bi)
        b.projection_init(gev.BI)
        b.projection_T0i_project(True, 1.0)
        t0i_ref = orc.projection_T0i_project(*args, phi=st["phi"])
        assert np.array_equal(b.download_field(gev.BI), t0i_ref)
        # plain CIC
        b.projection_init(gev.SOURCE)
        b.scalarProjectionCIC_project(1.0)
        cic_ref = orc.scalar_projection_cic_project(N, st["pos"], st["ids"], st["pmass"])
        assert np.array_equal(b.download_field(gev.SOURCE), cic_ref)
        # accumulation onto a non-zero target (projection without init)
        base = smooth_field(N, np.random.default_rng(9), 0.1)
        b.upload_field(gev.SOURCE, base)
        b.projection_T00_project(a, True, 1.0)
        acc_ref = orc.projection_T00_project(*args, a=a, phi=st["phi"], out=base.copy())
        assert np.array_equal(b.download_field(gev.SOURCE), acc_ref)


def test_deposit_edge_cases(gev, orc):
    N = 8
    # particles on cell faces, at 0 and just below 1, many in one cell, empty cells
    pos = np.array([[0.0, 0.0, 0.0], [np.nextafter(1.0, 0.0)] * 3, [0.5, 0.5, 0.5],
                    [0.5, 0.5, 0.5], [0.51, 0.52, 0.53], [0.125, 0.25, 0.375],
                    [0.999, 0.001, 0.5]] + [[0.3, 0.3, 0.3]] * 20)
    rng = np.random.default_rng(0)
    q = rng.standard_normal(pos.shape) * 1e-3
    ids = np.arange(len(pos), dtype=np.int64)[::-1].copy()  # IDs descending in storage
    phi = smooth_field(N, rng, 1e-5)
    with gev.Box(N, len(pos)) as b:
        b.upload_particles(pos, q, ids, 1e-3)
        b.upload_field(gev.PHI, phi)
        check_filing(gev, orc, b, N, pos, ids)
        b.projection_init(gev.SOURCE)
        b.projection_init(gev.SIJ)
        b.projection_T00_Tij_project(0.5, 1.0)
        args = (N, pos, q, ids, 1e-3)
        assert np.array_equal(b.download_field(gev.SOURCE), orc.projection_T00_project(*args, a=0.5, phi=phi))
        assert np.array_equal(b.download_field(gev.SIJ), orc.projection_Tij_project(*args, a=0.5, phi=phi))


def test_dense_cells_filing_and_deposit(gev, orc):
    """Cells of every size class of the filing's ID sort (tiny, one
    wavefront, one workgroup in LDS, and above the LDS size) with shuffled
    IDs: the filing equals the oracle's (ascending ID inside each cell), and
    the deposit, whose sums run in that order, is bit-exact -- before and
    after a drift re-files the particles."""
    N = 16
    rng = np.random.default_rng(5)
    groups = [(3, 0.41), (20, 0.27), (100, 0.66), (2500, 0.09), (1, 0.8)]
    pos = [np.tile([[c0, c0 + 0.01, c0 + 0.02]], (n, 1)) + rng.uniform(0, 0.05 / N, size=(n, 3))
           for n, c0 in groups]
    pos = np.concatenate(pos + [rng.uniform(0, 1, size=(300, 3))])
    n = len(pos)
    ids = rng.permutation(np.arange(10 * n, dtype=np.int64))[:n]
    q = rng.standard_normal((n, 3)) * 1e-3
    phi = smooth_field(N, rng, 1e-5)
    chi = smooth_field(N, rng, 1e-6)
    Bi = smooth_field(N, rng, 1e-8, ncomp=3)
    a = 0.3
    with gev.Box(N, n) as b:
        b.upload_particles(pos, q, ids, 1e-3)
        b.upload_field(gev.PHI, phi)
        b.upload_field(gev.CHI, chi)
        b.upload_field(gev.BI, Bi)
        for step in range(2):
            dpos, dq, dids = b.download_particles()
            check_filing(gev, orc, b, N, dpos, dids)
            b.projection_init(gev.SOURCE)
            b.projection_init(gev.SIJ)
            b.projection_T00_Tij_project(a, 1.0)
            args = (N, dpos, dq, dids, 1e-3)
            assert np.array_equal(b.download_field(gev.SOURCE), orc.projection_T00_project(*args, a=a, phi=phi))
            assert np.array_equal(b.download_field(gev.SIJ), orc.projection_Tij_project(*args, a=a, phi=phi))
            b.moveParticles(gev.POS_GR, 0.002, 3, [a, a * a * N])


@pytest.mark.parametrize("N", [64, 48])
def test_clustered_tiles_deposit_bit_exact(gev, orc, N):
    """Gaussian blobs a cell wide over a thin uniform background give cells
    of 1 to several hundred particles: tiles whose source region fits the
    LDS and tiles that take several chunks, interior tiles and tiles that
    wrap (a blob sits on the origin), and many source cells at or above the
    dense-cell size, whose partial sums are formed by (cell, target) lanes
    before the target threads add them.  The fused T00 + Tij deposit and
    Tij alone are bit-exact against the oracle."""
    rng = np.random.default_rng(N + 1)
    centres = np.concatenate([[[0.3 / N, 0.2 / N, 0.1 / N]], rng.uniform(0, 1, size=(40, 3))])
    sizes = rng.integers(20, 1500, size=len(centres))
    blobs = [c + rng.normal(0, 0.8 / N, size=(s, 3)) for c, s in zip(centres, sizes)]
    pos = np.mod(np.concatenate(blobs + [rng.uniform(0, 1, size=(N ** 3 // 8, 3))]), 1.0)
    pos[pos >= 1.0] = 0.0
    n = len(pos)
    ids = rng.permutation(n).astype(np.int64)
    q = rng.standard_normal((n, 3)) * 1e-3
    phi = smooth_field(N, rng, 1e-5)
    a = 0.7
    with gev.Box(N, n) as b:
        b.upload_particles(pos, q, ids, 1e-3)
        b.upload_field(gev.PHI, phi)
        counts, _ = b.download_filing()
        dpos, dq, dids = b.download_particles()
        args = (N, dpos, dq, dids, 1e-3)
        t00 = orc.projection_T00_project(*args, a=a, phi=phi)
        tij = orc.projection_Tij_project(*args, a=a, phi=phi)
        b.projection_init(gev.SOURCE)
        b.projection_init(gev.SIJ)
        b.projection_T00_Tij_project(a, 1.0)
        assert np.array_equal(b.download_field(gev.SOURCE), t00)
        assert np.array_equal(b.download_field(gev.SIJ), tij)
        b.projection_init(gev.SIJ)
        b.projection_Tij_project(a, True, 1.0)
        assert np.array_equal(b.download_field(gev.SIJ), tij)
        print("N=%d: %d particles, largest cell %d, %d cells of 5 or more" %
              (N, n, counts.max(), int(np.count_nonzero(counts >= 5))), flush=True)


def test_empty_box(gev):
    N = 16
    with gev.Box(N, 0) as b:
        b.projection_init(gev.SOURCE)
        b.projection_T00_project(0.1, True, 1.0)
        assert not b.download_field(gev.SOURCE).any()
        assert b.updateVel(gev.Q_GR, 0.01, 3, [0.1, 0.01 * N]) == 0.0
        b.moveParticles(gev.POS_GR, 0.01, 3, [0.1, 0.01 * N])
        counts, cells = b.download_filing()
        assert counts.sum() == 0 and cells.size == 0


def test_bad_arguments_are_errors(gev):
    with pytest.raises(gev.GevError):
        gev.Box(10, 5)  # 10 = 2*5 has no supported FFT
    with gev.Box(8, 2) as b:
        with pytest.raises(gev.GevError):
            b.upload_particles(np.zeros((3, 3)), np.zeros((3, 3)), np.arange(3), 1.0)  # > capacity
        with pytest.raises(gev.GevError):
            b.upload_particles(np.array([[1.0, 0.0, 0.0]]), np.zeros((1, 3)), np.arange(1), 1.0)
        with pytest.raises(gev.GevError):
            b.updateVel(7, 0.1, 3, [0.1, 1.0])


# ----------------------------------------------------------------- kicks / drifts

@pytest.mark.parametrize("nfield", [1, 2, 3])
def test_update_vel_gr_bit_exact(gev, orc, nfield):
    N = 24
    st = gr_state(N, 21)
    prm = [0.02, 0.02 * 0.02 * N]
    with new_box(gev, st, N) as b:
        mv = b.updateVel(gev.Q_GR, 0.013, nfield, prm)
        q_ref, mv_ref = orc.update_vel(N, st["pos"], st["q"], orc.Q_GR, 0.013, st["phi"], st["chi"],
                                       st["Bi"], nfield, prm)
        _, q, _ = b.download_particles()
        assert np.array_equal(q, q_ref)
        assert mv == mv_ref


def test_update_vel_newton_bit_exact(gev, orc):
    N = 16
    st = gr_state(N, 22)
    prm = [0.3, 0.3 * 0.3 * N]
    with new_box(gev, st, N) as b:
        mv = b.updateVel(gev.Q_NEWTON, 0.02, 1, prm)
        q_ref, mv_ref = orc.update_vel(N, st["pos"], st["q"], orc.Q_NEWTON, 0.02, st["phi"], st["chi"],
                                       None, 1, prm)
        assert np.array_equal(b.download_particles()[1], q_ref)
        assert mv == mv_ref


@pytest.mark.parametrize("method,nfield", [(0, 0), (0, 3), (1, 0)])
def test_move_particles_bit_exact(gev, orc, method, nfield):
    N = 24
    st = gr_state(N, 23)
    prm = [0.05, 0.05 * 0.05 * N]
    st["q"] *= 50.0  # move particles across cell and box boundaries
    with new_box(gev, st, N) as b:
        b.moveParticles(method, 0.02, nfield, prm)
        pos_ref = orc.move_particles(N, st["pos"], st["q"], method, 0.02, st["phi"], st["chi"], st["Bi"],
                                     nfield, prm)
        pos, _, ids = b.download_particles()
        # a drift stores the particles in cell order: match them by ID
        assert np.array_equal(pos[np.argsort(ids)], pos_ref[np.argsort(st["ids"])])
        assert np.all((pos >= 0) & (pos < 1))
        check_filing(gev, orc, b, N, pos_ref, st["ids"])


def test_kick_drift_fused_equals_separate(gev, orc):
    N = 16
    st = gr_state(N, 24)
    pk = [0.02, 0.02 ** 2 * N]
    pd = [0.021, 0.021 ** 2 * N]
    with new_box(gev, st, N) as b1, new_box(gev, st, N) as b2:
        m1 = b1.kick_drift(gev.Q_GR, 0.01, 3, pk, gev.POS_GR, 0.02, 3, pd)
        m2 = b2.updateVel(gev.Q_GR, 0.01, 3, pk)
        b2.moveParticles(gev.POS_GR, 0.02, 3, pd)
        p1, q1, _ = b1.download_particles()
        p2, q2, _ = b2.download_particles()
        assert m1 == m2 and np.array_equal(p1, p2) and np.array_equal(q1, q2)
        assert np.array_equal(b1.download_filing()[0], b2.download_filing()[0])


# ----------------------------------------------------------------- whole cycles

def run_both(gev, orc, N, boxsize, ncycles, gr=True, elliptic=False, seed=31, npc=1):
    st = gr_state(N, seed, npc_side=npc)
    sp = sim_params(N, boxsize, gr=gr, vector_elliptic=elliptic)
    cosmo_o = orc.make_cosmo()
    sim_o = orc.Sim(N, sp["gr_flag"], sp["vector_flag"], sp["Cf"], sp["steplimit"], sp["z_relax"],
                    sp["fourpiG"], cosmo_o)
    a0 = 1.0 / (1.0 + sp["z_in"])
    dt0 = orc.lib().orc_next_dtau(orc.ctypes.byref(sim_o), a0)
    so = orc.State(a0, 0.0, dt0, 0.0, 0.0, 0.0, 0.0, 0)
    ob = orc.OracleBox(N, len(st["ids"]))
    ob.pos[:] = st["pos"]; ob.q[:] = st["q"]; ob.id[:] = st["ids"]; ob.box.pmass = st["pmass"]
    ob.phi[:] = st["phi"]; ob.chi[:] = st["chi"]; ob.Bi[:] = st["Bi"]
    ob.BiFT[:] = st["BiFT"]; ob.scalarFT[:] = st["scalarFT"]
    ob.file()
    sim_g = gev.SimParams(sp["gr_flag"], sp["vector_flag"], sp["Cf"], sp["steplimit"], sp["z_relax"],
                          sp["fourpiG"], gev.make_cosmology())
    sg = gev.CycleState(a0, 0.0, dt0, 0.0, 0.0, 0.0, 0.0, 0, 0)
    b = new_box(gev, st, N)
    out = []
    for c in range(ncycles):
        ob.cycle(sim_o, so)
        b.run_cycles(sim_g, sg, 1)
        out.append((so.a, sg.a, so.dtau, sg.dtau))
        assert so.a == sg.a and so.dtau == sg.dtau and so.tau == sg.tau  # host scalars bit-exact
        pos, q, ids = b.download_particles()
        order = np.argsort(ids)
        for name, f in (("phi", gev.PHI), ("chi", gev.CHI), ("Bi", gev.BI)):
            g = b.download_field(f)
            r = getattr(ob, name)
            assert rel(g, r) < 1e-10, (c, name, rel(g, r))
        assert rel(pos[order], ob.pos) < 1e-10
        assert rel(q[order], ob.q) < 1e-10
        counts, cells = b.download_filing()
        assert np.array_equal(counts, ob.count)
        assert np.array_equal(cells[order], ob.cell)
        assert abs(sg.T00hom - so.T00hom) <= 1e-12 * abs(so.T00hom) or not gr
        assert abs(sg.maxvel - so.maxvel) <= 1e-10 * abs(so.maxvel)
        assert abs(sg.phi_k0 - so.phi_k0) <= 1e-10 * abs(so.phi_k0) + 1e-300
        assert rel(b.download_field(gev.BIFT), ob.BiFT) < 1e-10, (c, "BiFT")
    b.close()
    ob.close()
    return out


def cycle_by_calls(gev, b, sp, sim_g, st):
    """One cycle through the unfused entry points, in main.cpp's order and
    with its coefficients, host scalars advanced as gev_run_cycles does."""
    N = b.N
    dx = 1.0 / N
    fpG = sp["fourpiG"]
    cs = sim_g.cosmo
    a = st["a"]
    gr = sp["gr_flag"] > 0
    b.projection_init(gev.SOURCE)
    b.projection_init(gev.SIJ)
    if gr:
        b.projection_T00_project(a, True, 1.0)
    else:
        b.scalarProjectionCIC_project(1.0)
    b.projection_Tij_project(a, True, 1.0)
    if sp["vector_flag"] == 1:
        b.projection_init(gev.BI)
        b.projection_T0i_project(True, 1.0)
    if gr:
        if st["dtau_old"] > 0.:
            H = gev.hconf(a, fpG, cs)
            b.prepareFTsource_scalar(cs.Omega_cdm + cs.Omega_b, 3. * H * dx * dx / st["dtau_old"],
                                     fpG * dx * dx / a, 3. * H * H * dx * dx)
            b.fft(gev.PLAN_SOURCE, gev.FFT_FORWARD)
            b.solveModifiedPoissonFT(1. / (dx * dx), 3. * gev.hconf(a, fpG, cs) / st["dtau_old"])
            b.fft(gev.PLAN_PHI, gev.FFT_BACKWARD)
    else:
        b.fft(gev.PLAN_SOURCE, gev.FFT_FORWARD)
        b.solveModifiedPoissonFT(fpG / a, 0.)
        b.fft(gev.PLAN_PHI, gev.FFT_BACKWARD)
    b.prepareFTsource_tensor(2. * fpG * dx * dx / a)
    b.fft(gev.PLAN_SIJ, gev.FFT_FORWARD)
    b.projectFTscalar(0)
    b.fft(gev.PLAN_CHI, gev.FFT_BACKWARD)
    if sp["vector_flag"] == 1:
        b.fft(gev.PLAN_BI, gev.FFT_FORWARD)
        b.projectFTvector(fpG * dx * dx, 0.)
    else:
        b.evolveFTvector(a * a * st["dtau_old"])
    if gr:
        b.fft(gev.PLAN_BI, gev.FFT_BACKWARD)
    nfk = (3 if 1. / a < sp["z_relax"] + 1. else 2) if gr else 1
    b.updateVel(gev.Q_GR if gr else gev.Q_NEWTON, (st["dtau"] + st["dtau_old"]) / 2., nfk,
                [a, a * a * N], want_maxvel=False)
    a = gev.rungekutta4bg(a, fpG, cs, 0.5 * st["dtau"])
    nfd = (3 if 1. / a < sp["z_relax"] + 1. else 0) if gr else 0
    b.moveParticles(gev.POS_GR if gr else gev.POS_NEWTON, st["dtau"], nfd, [a, a * a * N])
    a = gev.rungekutta4bg(a, fpG, cs, 0.5 * st["dtau"])
    st["tau"] += st["dtau"]
    st["dtau_old"] = st["dtau"]
    st["a"] = a
    st["dtau"] = gev.next_dtau(sim_g, N, a)


@pytest.mark.parametrize("N,gr,elliptic", [(32, True, False), (48, True, False), (32, True, True),
                                           (32, False, False), (16, True, False)])
def test_fused_cycle_equals_call_sequence(gev, N, gr, elliptic):
    """gev_run_cycles runs the k-space solves inside the FFT z passes and the
    kick and drift in one particle pass; the unfused entry points, called in
    main.cpp's order, must give the same bits for every carried field."""
    st = gr_state(N, 5)
    sp = sim_params(N, 5.0 * N, gr=gr, vector_elliptic=elliptic)
    sim_g = gev.SimParams(sp["gr_flag"], sp["vector_flag"], sp["Cf"], sp["steplimit"], sp["z_relax"],
                          sp["fourpiG"], gev.make_cosmology())
    a0 = 1.0 / (1.0 + sp["z_in"])
    dt0 = gev.next_dtau(sim_g, N, a0)
    fused, calls = new_box(gev, st, N), new_box(gev, st, N)
    sg = gev.CycleState(a0, 0.0, dt0, 0.0, 0.0, 0.0, 0.0, 0, 0)
    sc = {"a": a0, "tau": 0.0, "dtau": dt0, "dtau_old": 0.0}
    for cyc in range(3):
        fused.run_cycles(sim_g, sg, 1)
        cycle_by_calls(gev, calls, sp, sim_g, sc)
        assert sg.a == sc["a"] and sg.dtau == sc["dtau"]
        for f in (gev.PHI, gev.CHI, gev.BI, gev.BIFT):
            x, y = fused.download_field(f), calls.download_field(f)
            assert np.array_equal(x, y), (cyc, f, rel(x, y))
        pf, qf, idf = fused.download_particles()
        pc, qc, idc = calls.download_particles()
        assert np.array_equal(pf, pc) and np.array_equal(qf, qc) and np.array_equal(idf, idc), cyc
    fused.close()
    calls.close()


def test_cycles_gr_parabolic_32(gev, orc):
    run_both(gev, orc, 32, 160.0, 4)


def test_cycles_gr_parabolic_48_radix3(gev, orc):
    run_both(gev, orc, 48, 240.0, 3)


def test_cycles_gr_elliptic_32(gev, orc):
    run_both(gev, orc, 32, 160.0, 3, elliptic=True)


def test_cycles_newton_32(gev, orc):
    run_both(gev, orc, 32, 160.0, 3, gr=False)


def test_cycles_gr_64_dense(gev, orc):
    run_both(gev, orc, 64, 320.0, 2, npc=1)


def test_cycles_gr_16_eight_per_cell(gev, orc):
    run_both(gev, orc, 16, 80.0, 3, npc=2)


# ----------------------------------------------------------------- full size

def test_full_size_512_properties(gev):
    """At the 512^3 workload size: FFT round trip, filing totals, mass
    conservation of the deposit, and bit-reproducibility of a cycle."""
    N = 512
    rng = np.random.default_rng(7)
    with gev.Box(N, N ** 3) as b:
        f = rng.standard_normal((N, N, N))
        b.upload_field(gev.SOURCE, f)
        b.fft(gev.PLAN_SOURCE, gev.FFT_FORWARD)
        b.fft(gev.PLAN_PHI, gev.FFT_BACKWARD)
        assert rel(b.download_field(gev.PHI), f * N ** 3) < 1e-12
        # a 512^3 particle lattice; plain CIC deposit conserves mass exactly up to rounding
        g = (np.arange(N) + 0.5) / N
        pos = np.empty((N ** 3, 3))
        zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")
        pos[:, 0] = xx.ravel(); pos[:, 1] = yy.ravel(); pos[:, 2] = zz.ravel()
        del xx, yy, zz
        pos += rng.uniform(-0.4, 0.4, pos.shape) / N
        np.mod(pos, 1.0, out=pos)
        q = rng.standard_normal(pos.shape) * 1e-3
        ids = np.arange(N ** 3, dtype=np.int64)
        b.upload_particles(pos, q, ids, 0.31 / N ** 3)
        counts, _ = b.download_filing()
        assert counts.sum() == N ** 3
        b.projection_init(gev.SOURCE)
        b.scalarProjectionCIC_project(1.0)
        assert abs(b.field_sum(gev.SOURCE) / N ** 3 - 0.31) < 1e-12
        # determinism: same cycle twice gives identical bits
        b.projection_init(gev.SOURCE)
        b.projection_init(gev.SIJ)
        b.projection_T00_Tij_project(0.01, 1.0)
        s1 = b.download_field(gev.SIJ)
        b.projection_init(gev.SIJ)
        b.projection_T00_Tij_project(0.01, 1.0)
        assert np.array_equal(b.download_field(gev.SIJ), s1)
