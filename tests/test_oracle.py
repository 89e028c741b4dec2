"""Pins the CPU oracle against facts independent of its own formulas:
numpy's FFT, eigenfunctions of the discrete operators built from real-space
stencils, continuum limits, closed forms of single particles and uniform
fields, conservation laws, and thread-count invariance.  A mistake such as a
dropped term, a wrong sign or index, or a transposed operand in the oracle
fails one of these.
"""
import numpy as np
import pytest

from states import gr_state, smooth_field

AX = (-3, -2, -1)


def rel(a, b):
    d = np.max(np.abs(np.asarray(a) - np.asarray(b)))
    s = np.max(np.abs(b))
    return d / s if s > 0 else d


def lap(f, N):
    """7-point Laplacian with dx = 1/N on a periodic (z,y,x) array."""
    out = -6.0 * f
    for ax in (-1, -2, -3):
        out = out + np.roll(f, 1, ax) + np.roll(f, -1, ax)
    return out * N * N


def kappa(N, m):
    """Backward difference symbol: N (1 - exp(-i 2 pi m/N)) = i kappa."""
    return -1j * N * (1 - np.exp(-2j * np.pi * m / N))


def kappa3(N):
    H = N // 2 + 1
    KZ, KY, KX = np.meshgrid(kappa(N, np.arange(N)), kappa(N, np.arange(N)), kappa(N, np.arange(H)),
                             indexing="ij")
    return [KX, KY, KZ]


# ------------------------------------------------------------------ FFT

@pytest.mark.parametrize("N", [4, 6, 8, 12, 16, 24, 48])
def test_fft_matches_numpy(orc, N):
    rng = np.random.default_rng(N)
    f = rng.standard_normal((N, N, N))
    assert rel(orc.fft_r2c(f), np.fft.rfftn(f)) < 1e-14
    k = rng.standard_normal((N, N, N // 2 + 1)) + 1j * rng.standard_normal((N, N, N // 2 + 1))
    assert rel(orc.fft_c2r(k), np.fft.irfftn(k, s=(N, N, N), axes=AX) * N ** 3) < 1e-14
    assert rel(orc.fft_c2r(orc.fft_r2c(f)), f * N ** 3) < 1e-14


@pytest.mark.parametrize("n", [2, 3, 8, 9, 12, 48, 96, 512])
def test_fft1d_matches_numpy(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert rel(orc.fft1d(x, -1), np.fft.fft(x)) < 1e-14
    assert rel(orc.fft1d(x, +1), np.fft.ifft(x) * n) < 1e-14


def test_fft_delta_and_single_mode(orc):
    N = 16
    f = np.zeros((N, N, N)); f[0, 0, 0] = 1.0
    assert np.allclose(orc.fft_r2c(f), 1.0, atol=0, rtol=1e-15)
    z, y, x = np.meshgrid(*[np.arange(N)] * 3, indexing="ij")
    f = np.cos(2 * np.pi * (3 * x + 2 * y - 5 * z) / N)
    k = orc.fft_r2c(f)
    assert abs(k[-5 % N, 2, 3] - N ** 3 / 2) < 1e-9
    k[-5 % N, 2, 3] = 0
    assert np.abs(k).max() < 1e-9


# ------------------------------------------------------------------ k space

@pytest.mark.parametrize("N,modif", [(16, 0.0), (12, 3.7)])
def test_poisson_solves_discrete_laplacian(orc, N, modif):
    """pot = c2r(solve(r2c(src))) satisfies (lap - modif) pot = coeff * src
    with the real-space 7-point Laplacian (mean removed when modif = 0)."""
    rng = np.random.default_rng(4)
    src = smooth_field(N, rng, 1.0)
    coeff = 0.37
    pot = orc.fft_c2r(orc.solve_modified_poisson_ft(orc.fft_r2c(src), coeff, modif))
    target = coeff * (src - src.mean()) if modif == 0 else coeff * src
    assert rel(lap(pot, N) - modif * pot, target) < 1e-11
    if modif == 0:
        assert abs(pot.mean()) < 1e-14 * np.abs(pot).max()


def _sym(S, i, j):
    idx = {(0, 0): 0, (0, 1): 1, (0, 2): 2, (1, 1): 3, (1, 2): 4, (2, 2): 5}
    return S[idx[(min(i, j), max(i, j))]]


def test_project_scalar_inverts_longitudinal_tensor(orc):
    """S_ij built from psi by real-space second differences (vertex diagonals,
    plaquette off-diagonals from forward differences) projects to chi = -psi/N
    in FFT units; a pure trace projects to zero."""
    N = 16
    rng = np.random.default_rng(6)
    psi = smooth_field(N, rng, 1.0, kcut=4)
    psi -= psi.mean()
    S = np.zeros((6, N, N, N))
    axes = {0: -1, 1: -2, 2: -3}
    fwd = lambda f, i: (np.roll(f, -1, axes[i]) - f) * N
    for i in range(3):
        _sym(S, i, i)[:] = -(np.roll(psi, -1, axes[i]) + np.roll(psi, 1, axes[i]) - 2 * psi) * N * N
    for (i, j) in ((0, 1), (0, 2), (1, 2)):
        _sym(S, i, j)[:] = -fwd(fwd(psi, i), j)
    SFT = np.fft.rfftn(S, axes=AX)
    chi = orc.project_ft_scalar(SFT)
    expect = -np.fft.rfftn(psi) / N
    expect[0, 0, 0] = 0
    assert rel(chi, expect) < 1e-12
    tr = np.zeros_like(SFT)
    tr[0] = tr[3] = tr[5] = np.fft.rfftn(psi)
    assert np.abs(orc.project_ft_scalar(tr)).max() < 1e-12 * np.abs(tr).max()
    # add = 1 accumulates
    chi2 = orc.project_ft_scalar(SFT, chiFT=chi, add=1)
    assert rel(chi2, 2 * expect) < 1e-12


def test_evolve_vector_is_transverse_spin1_projection(orc):
    N = 12
    rng = np.random.default_rng(7)
    H = N // 2 + 1
    kap = kappa3(N)
    S = rng.standard_normal((6, N, N, H)) + 1j * rng.standard_normal((6, N, N, H))
    zero = np.zeros((3, N, N, H), complex)
    dB = orc.evolve_ft_vector(S, zero, 0.7)
    assert np.abs(sum(kap[i] * dB[i] for i in range(3))).max() < 1e-13 * np.abs(dB).max()
    # a scalar-derived tensor has no spin-1 part; the trace neither
    psi = rng.standard_normal((N, N, H)) + 1j * rng.standard_normal((N, N, H))
    Ss = np.zeros_like(S)
    for i in range(3):
        for j in range(i, 3):
            _sym(Ss, i, j)[:] = np.abs(kap[i]) ** 2 * psi if i == j else np.conj(kap[i] * kap[j]) * psi
    assert np.abs(orc.evolve_ft_vector(Ss, zero, 0.7)).max() < 1e-12 * np.abs(dB).max()
    # linear in a2dtau and additive onto B; k = 0 is zeroed
    B0 = rng.standard_normal((3, N, N, H)) + 1j * rng.standard_normal((3, N, N, H))
    out = orc.evolve_ft_vector(S, B0, 1.4)
    expect = B0 + 2 * dB
    expect[:, 0, 0, 0] = 0
    assert rel(out, expect) < 1e-13


def test_project_vector_transverse_and_elliptic(orc):
    N = 12
    rng = np.random.default_rng(8)
    H = N // 2 + 1
    kap = kappa3(N)
    V = rng.standard_normal((3, N, N, H)) + 1j * rng.standard_normal((3, N, N, H))
    B = orc.project_ft_vector(V, 0.3, 0.2)
    assert np.abs(sum(kap[i] * B[i] for i in range(3))).max() < 1e-13 * np.abs(B).max()
    # gradients (S_i = conj(kappa_i) psi) are removed
    psi = V[0]
    G = np.stack([np.conj(k) * psi for k in kap])
    assert np.abs(orc.project_ft_vector(G, 0.3, 0.2)).max() < 1e-13 * np.abs(G).max()
    # a transverse input (curl with backward differences) solves (k^2+modif) B = 4 coeff S
    A = V
    curl = np.stack([kap[1] * A[2] - kap[2] * A[1], kap[2] * A[0] - kap[0] * A[2],
                     kap[0] * A[1] - kap[1] * A[0]])
    Bc = orc.project_ft_vector(curl, 0.3, 0.2)
    k2 = sum(np.abs(k) ** 2 for k in kap)
    expect = curl * 4 * 0.3 / (k2 + 0.2)
    expect[:, 0, 0, 0] = 0
    assert rel(Bc, expect) < 1e-13


# ------------------------------------------------------------------ stencils

def test_prepare_tensor_continuum_limit_and_exact_cases(orc):
    N = 64
    z, y, x = np.meshgrid(*[np.arange(N)] * 3, indexing="ij")
    A, m = 1e-3, (1, 2, 1)
    th = 2 * np.pi * (m[0] * x + m[1] * y + m[2] * z) / N
    phi = A * np.sin(th)
    T = np.random.default_rng(1).standard_normal((6, N, N, N))
    S = orc.prepare_ft_source_tensor(phi, T, 0.25)
    # continuum: S_ij = coeff T_ij + 2 dx^2 d_i phi d_j phi, with d_i phi = A k_i cos(th),
    # evaluated where S_ij lives: vertices (i = j), plaquette centres x + (e_i + e_j)/2
    kphys = [2 * np.pi * mi for mi in m]  # dx = 1/N, d/dx in box units
    dx = 1.0 / N
    for i in range(3):
        for j in range(i, 3):
            shift = 0.0 if i == j else np.pi * (m[i] + m[j]) / N
            cont = 0.25 * _sym(T, i, j) + 2 * dx * dx * A * A * kphys[i] * kphys[j] * np.cos(th + shift) ** 2
            err = np.abs(_sym(S, i, j) - cont).max()
            assert err < 0.05 * 2 * dx * dx * A * A * kphys[i] * kphys[j] + 1e-15
    # constant phi: only the coeff * T term survives (exactly)
    S0 = orc.prepare_ft_source_tensor(np.full((N, N, N), 0.3), T, 0.25)
    assert np.abs(S0 - 0.25 * T).max() < 1e-15


def test_prepare_scalar_continuum_limit_and_exact_cases(orc):
    N = 64
    z, y, x = np.meshgrid(*[np.arange(N)] * 3, indexing="ij")
    A = 1e-3
    th = 2 * np.pi * (x + 2 * y + z) / N
    phi = A * np.sin(th)
    rng = np.random.default_rng(2)
    chi = rng.standard_normal((N, N, N)) * 1e-4
    src = 0.31 + rng.standard_normal((N, N, N)) * 0.1
    c1, c2, c3 = 0.3, 0.02, 0.7
    r = orc.prepare_ft_source_scalar(phi, chi, src, 0.31, c1, c2, c3)
    dx = 1.0 / N
    grad2 = (A * 2 * np.pi) ** 2 * (1 + 4 + 1) * np.cos(th) ** 2
    cont = c2 * (src - 0.31) * (1 - 2 * phi) + 0.5 * dx * dx * grad2 + (c3 - c1) * phi - c3 * chi
    assert np.abs(r - cont).max() < 0.02 * 0.5 * dx * dx * (A * 2 * np.pi) ** 2 * 6 + 1e-16
    r0 = orc.prepare_ft_source_scalar(np.zeros((N, N, N)), chi, src, 0.31, c1, c2, c3)
    assert np.abs(r0 - (c2 * (src - 0.31) - c3 * chi)).max() < 1e-16


# ------------------------------------------------------------------ filing

def test_filing_counts_and_order(orc):
    N = 8
    rng = np.random.default_rng(3)
    pos = rng.uniform(0, 1, (500, 3))
    pos[:5] = [[0, 0, 0], [0.125, 0.125, 0.125], [np.nextafter(1, 0)] * 3, [0.5, 0.25, 0.75], [0.999, 0.0, 0.4]]
    ids = rng.permutation(500).astype(np.int64)
    cell, count, start, perm = orc.file_particles(N, pos, ids)
    c = np.floor(pos * N).astype(int)  # N a power of two: pos/dx == pos*N exactly
    assert np.array_equal(cell, c[:, 0] + N * (c[:, 1] + N * c[:, 2]))
    assert count.sum() == 500 and np.array_equal(count, np.bincount(cell, minlength=N ** 3))
    assert start[-1] == 500 and np.array_equal(np.diff(start), count)
    assert np.array_equal(cell[perm], np.sort(cell))
    for s in range(N ** 3):
        idl = ids[perm[start[s]:start[s + 1]]]
        assert np.all(np.diff(idl) > 0)


# ------------------------------------------------------------------ deposits

def one_particle(N, p, qv=(0, 0, 0)):
    return np.array([p], dtype=float), np.array([qv], dtype=float), np.array([0], dtype=np.int64)


def test_t00_single_particle_closed_form(orc):
    N, a, m = 8, 0.1, 0.01
    pos, q, ids = one_particle(N, (0.30, 0.55, 0.95), (1e-3, -2e-3, 5e-4))
    phi = np.zeros((N, N, N))
    T = orc.projection_T00_project(N, pos, q, ids, m, a=a, phi=phi)
    c = np.floor(pos[0] * N).astype(int)
    w = pos[0] * N - c
    q2 = (q[0] ** 2).sum()
    e = np.sqrt(q2 + a * a)
    mass = m * N ** 3 / a
    for dz in (0, 1):
        for dy in (0, 1):
            for dxx in (0, 1):
                W = (w[0] if dxx else 1 - w[0]) * (w[1] if dy else 1 - w[1]) * (w[2] if dz else 1 - w[2])
                v = T[(c[2] + dz) % N, (c[1] + dy) % N, (c[0] + dxx) % N]
                assert abs(v - W * e * mass) < 1e-13 * mass
    assert abs(T.sum() - e * mass) < 1e-13 * mass
    # with a uniform phi0 the energy density picks up the factor (e + (3e + q^2/e) phi0)
    T2 = orc.projection_T00_project(N, pos, q, ids, m, a=a, phi=np.full((N, N, N), 1e-3))
    assert abs(T2.sum() - (e + (3 * e + q2 / e) * 1e-3) * mass) < 1e-12 * mass


def test_projections_conserve_totals(orc):
    N = 16
    st = gr_state(N, 5)
    pos, q, ids, m = st["pos"], st["q"], st["ids"], st["pmass"]
    a = 0.05
    zero = np.zeros((N, N, N))
    # CIC mass: every particle deposits coeff * m / dx^3 in total
    rho = orc.scalar_projection_cic_project(N, pos, ids, m, coeff=2.0)
    assert abs(rho.sum() - 2.0 * m * N ** 3 * len(ids)) < 1e-12 * rho.sum()
    # T00 without phi: sum_p e_p m / (dx^3 a)
    e = np.sqrt((q ** 2).sum(1) + a * a)
    T = orc.projection_T00_project(N, pos, q, ids, m, a=a, phi=zero)
    assert abs(T.sum() - (e * m * N ** 3 / a).sum()) < 1e-12 * T.sum()
    # Tij: sum = sum_p m q_i q_j / (e a dx^3) (phi = 0)
    Tij = orc.projection_Tij_project(N, pos, q, ids, m, a=a, phi=zero)
    for i in range(3):
        for j in range(i, 3):
            ref = (m * q[:, i] * q[:, j] / e).sum() * N ** 3 / a
            assert abs(_sym(Tij, i, j).sum() - ref) < 1e-10 * abs(ref) + 1e-30
    # T0i: sum = sum_p m q_i / dx^3
    T0i = orc.projection_T0i_project(N, pos, q, ids, m, phi=zero)
    for i in range(3):
        ref = (m * q[:, i]).sum() * N ** 3
        assert abs(T0i[i].sum() - ref) < 1e-9 * np.abs(m * q[:, i]).sum() * N ** 3
    # at rest there is no stress
    assert not orc.projection_Tij_project(N, pos, np.zeros_like(q), ids, m, a=a, phi=st["phi"]).any()


def test_projection_on_lattice_corners_is_uniform(orc):
    N = 8
    g = np.arange(N) / N
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    q = np.zeros_like(pos)
    ids = np.arange(N ** 3, dtype=np.int64)
    m = 0.31 / N ** 3
    T = orc.projection_T00_project(N, pos, q, ids, m, a=0.2, phi=np.zeros((N, N, N)))
    assert np.allclose(T, 0.31, rtol=1e-15, atol=0)


def test_projection_shift_invariance(orc):
    N = 16
    st = gr_state(N, 6)
    pos = st["pos"]
    shifted = np.mod(pos + np.array([1.0 / N, 2.0 / N, 0.0]), 1.0)
    args = dict(a=0.03, phi=st["phi"])
    T1 = orc.projection_Tij_project(N, pos, st["q"], st["ids"], st["pmass"], **args)
    args2 = dict(a=0.03, phi=np.roll(st["phi"], (2, 1), axis=(-2, -1)))
    T2 = orc.projection_Tij_project(N, shifted, st["q"], st["ids"], st["pmass"], **args2)
    assert rel(T2, np.roll(T1, (2, 1), axis=(-2, -1))) < 1e-12


def test_projection_independent_of_thread_count(orc):
    """The ordered scatter adds onto every target in sequential site order for
    any thread count, so results are bit-identical."""
    N = 16
    st = gr_state(N, 7, npc_side=2)
    outs = []
    for nt in (1, 3, 8):
        orc.set_threads(nt)
        outs.append((orc.projection_T00_project(N, st["pos"], st["q"], st["ids"], st["pmass"], a=0.04, phi=st["phi"]),
                     orc.projection_Tij_project(N, st["pos"], st["q"], st["ids"], st["pmass"], a=0.04, phi=st["phi"]),
                     orc.projection_T0i_project(N, st["pos"], st["q"], st["ids"], st["pmass"], phi=st["phi"])))
    orc.set_threads(8)
    for o in outs[1:]:
        for x, y in zip(o, outs[0]):
            assert np.array_equal(x, y)


# ------------------------------------------------------------------ kicks / drifts

def test_kick_uniform_fields_leave_q_unchanged(orc):
    N = 8
    st = gr_state(N, 8)
    phi = np.full((N, N, N), 1e-4)
    chi = np.full((N, N, N), -3e-5)
    B = np.full((3, N, N, N), 2e-6)
    prm = [0.1, 0.1 * 0.1 * N]
    qn, mv = orc.update_vel(N, st["pos"], st["q"], orc.Q_GR, 0.05, phi, chi, B, 3, prm)
    assert np.abs(qn - st["q"]).max() < 1e-18
    assert abs(mv - np.sqrt(((st["q"] ** 2).sum(1)).max()) / 0.1) < 1e-12 * mv


def test_kick_linear_potential_closed_form(orc):
    """phi = alpha * (x index) near the particle gives a CIC gradient alpha per
    cell: q_x -= dtau e alpha (v^2 + e^2) / e^2 / dx (GR), dtau a alpha / dx (Newton);
    B_y = beta * (x index) gives the frame-dragging term q_y beta / (a^2 N e)."""
    N = 16
    z, y, x = np.meshgrid(*[np.arange(N, dtype=float)] * 3, indexing="ij")
    alpha, beta = 1e-5, 3e-6
    phi = alpha * x
    chi = np.zeros((N, N, N))
    B = np.zeros((3, N, N, N))
    B[1] = beta * x
    pos = np.array([[0.41, 0.52, 0.33], [0.27, 0.74, 0.61]])
    q = np.array([[1e-3, 2e-3, -1e-3], [-2e-3, 1e-3, 3e-3]])
    a, dtau, dx = 0.2, 0.01, 1.0 / N
    prm = [a, a * a * N]
    for nfield in (1, 3):
        qn, _ = orc.update_vel(N, pos, q, orc.Q_GR, dtau, phi, chi, B, nfield, prm)
        v2 = (q ** 2).sum(1)
        e2 = v2 + a * a
        e = np.sqrt(e2)
        g = alpha * (v2 + e2) / e2
        if nfield >= 3:
            g = g + q[:, 1] * beta / prm[1] / e
        assert np.allclose(qn[:, 0], q[:, 0] - dtau * e * g / dx, rtol=0, atol=1e-17)
        assert np.allclose(qn[:, 1:], q[:, 1:], rtol=0, atol=1e-18)
    qn, _ = orc.update_vel(N, pos, q, orc.Q_NEWTON, dtau, phi, chi, None, 1, prm)
    assert np.allclose(qn[:, 0], q[:, 0] - dtau * a * alpha / dx, rtol=0, atol=1e-17)


def test_drift_closed_forms(orc):
    N = 16
    st = gr_state(N, 9)
    pos, q = st["pos"], st["q"] * 100
    a, dtau = 0.05, 0.03
    prm = [a, a * a * N]
    zero = np.zeros((N, N, N))
    p0 = orc.move_particles(N, pos, q, orc.POS_GR, dtau, zero, zero, np.zeros((3, N, N, N)), 0, prm)
    e = np.sqrt((q ** 2).sum(1, keepdims=True) + a * a)
    assert np.abs(np.mod(pos + dtau * q / e, 1.0) - p0).max() < 1e-15
    pn = orc.move_particles(N, pos, q, orc.POS_NEWTON, dtau, None, None, None, 0, prm)
    assert np.abs(np.mod(pos + dtau * q / a, 1.0) - pn).max() < 1e-15
    # uniform phi0, chi0, B0: v = q (1 + (3 - v^2/e^2) phi0 - chi0)/e, plus B0/(a^2 N)
    ph0, ch0, b0 = 1e-3, 2e-4, 5e-6
    p3 = orc.move_particles(N, pos, q, orc.POS_GR, dtau, np.full((N, N, N), ph0), np.full((N, N, N), ch0),
                            np.full((3, N, N, N), b0), 3, prm)
    v2 = (q ** 2).sum(1, keepdims=True)
    v = q * (1 + (3 - v2 / e ** 2) * ph0 - ch0) / e
    assert np.abs(np.mod(pos + dtau * (v + b0 / prm[1]), 1.0) - p3).max() < 1e-15


def test_drift_wraps_into_unit_box(orc):
    N = 8
    pos = np.array([[0.99, 0.01, 0.5], [0.0, 0.999999, 0.5]])
    q = np.array([[1.0, -1.0, 0.0], [-1.0, 1.0, 0.0]]) * 0.5
    p = orc.move_particles(N, pos, q, orc.POS_NEWTON, 0.05, None, None, None, 0, [1.0, N])
    assert np.all((p >= 0) & (p < 1))
    assert np.allclose(p[0], [0.015, 0.985, 0.5]) and np.allclose(p[1], [0.975, 0.024999, 0.5])


# ------------------------------------------------------------------ background

def test_background_eds_closed_form(orc):
    """Omega_m = 1: a(tau) = (fourpiG/6) tau^2, H = 2/tau, tau = sqrt(6 a / fourpiG)."""
    c = orc.Cosmo(1.0, 0.0, 1.0, 0.0, 0.0, 0.0, -1.0, 0.0, 0.7)
    fpg = 0.5
    a0 = 0.01
    tau0 = np.sqrt(6 * a0 / fpg)
    assert abs(orc.hconf(a0, fpg, c) - 2 / tau0) < 1e-14 * 2 / tau0
    ph = orc.particle_horizon(a0, fpg, c)
    assert abs(ph - tau0 * (1 - 1e-7)) < 1e-12 * tau0
    a, tau, dt = a0, tau0, tau0 / 200
    for _ in range(400):
        a = orc.rk4bg(a, fpg, c, dt)
        tau += dt
    assert abs(a - fpg / 6 * tau * tau) < 1e-9 * a


def test_background_lcdm_against_scipy(orc):
    from scipy.integrate import quad, solve_ivp
    c = orc.make_cosmo()
    fpg = 1.5 * 320.0 ** 2 / 2997.92458 ** 2
    om, ol = 0.31, 0.69
    H = lambda a: np.sqrt(2 * fpg / 3 * (om / a + ol * a * a))
    assert abs(orc.hconf(0.3, fpg, c) - H(0.3)) < 1e-14 * H(0.3)
    sol = solve_ivp(lambda t, y: [y[0] * H(y[0])], (0, 2.0), [0.01], rtol=1e-12, atol=1e-15)
    a, dt = 0.01, 2.0 / 1000
    for _ in range(1000):
        a = orc.rk4bg(a, fpg, c, dt)
    assert abs(a - sol.y[0, -1]) < 1e-9 * a
    tau = quad(lambda s: 2 / (s * np.sqrt(2 / 3 * (om / s ** 2 + ol * s ** 4))), np.sqrt(0.01) * 1e-7,
               np.sqrt(0.01), epsabs=0, epsrel=1e-13)[0] / np.sqrt(fpg)
    assert abs(orc.particle_horizon(0.01, fpg, c) - tau) < 1e-11 * tau


# ------------------------------------------------------------------ cycle

def test_homogeneous_box_stays_homogeneous(orc):
    """Particles on cell corners at rest with phi = chi = B = 0: T00 equals the
    background, every source vanishes and nothing moves over several cycles."""
    N = 8
    g = np.arange(N) / N
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    ob = orc.OracleBox(N, N ** 3)
    ob.pos[:] = pos; ob.q[:] = 0; ob.id[:] = np.arange(N ** 3); ob.box.pmass = 0.31 / N ** 3
    ob.file()
    sim = orc.Sim(N, 1, 0, 48.0, 0.04, 100.0, 1.5 * 40 ** 2 / 2997.92458 ** 2, orc.make_cosmo())
    a0 = 1 / 101.
    st = orc.State(a0, 0, orc.lib().orc_next_dtau(orc.ctypes.byref(sim), a0), 0, 0, 0, 0, 0)
    for _ in range(4):
        ob.cycle(sim, st)
    assert abs(st.T00hom - 0.31) < 1e-14
    assert np.abs(ob.phi).max() < 1e-17 and np.abs(ob.chi).max() < 1e-17 and np.abs(ob.Bi).max() < 1e-17
    assert np.array_equal(ob.pos, pos) and not ob.q.any()
    ob.close()


def test_cycle_newton_linear_growth(orc):
    """Newtonian cycles of a small single-mode displacement in an Einstein-de
    Sitter background grow the displacement like a (growing mode D = a),
    to a few per cent over a factor 2 in a."""
    N = 16
    Lbox = 1000.0
    fpg = 1.5 * Lbox ** 2 / 2997.92458 ** 2
    c = orc.Cosmo(1.0, 0.0, 1.0, 0.0, 0.0, 0.0, -1.0, 0.0, 0.7)
    g = (np.arange(N) + 0.5) / N
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    A = 0.02 / N
    a0 = 0.02
    psi = A * np.sin(2 * np.pi * x)
    H0 = np.sqrt(2 * fpg / 3 / a0)
    pos = np.stack([np.mod(x + psi, 1.0).ravel(), y.ravel(), z.ravel()], 1)
    q = np.zeros_like(pos)
    q[:, 0] = (a0 * H0 * psi).ravel()  # q = a v, v = dpsi/dtau = H psi for D = a
    ob = orc.OracleBox(N, N ** 3)
    ob.pos[:] = pos; ob.q[:] = q; ob.id[:] = np.arange(N ** 3); ob.box.pmass = 1.0 / N ** 3
    ob.file()
    sim = orc.Sim(N, 0, 0, 48.0, 0.01, 1 / a0 - 1, fpg, c)
    st = orc.State(a0, 0, orc.lib().orc_next_dtau(orc.ctypes.byref(sim), a0), 0, 0, 0, 0, 0)
    while st.a < 2 * a0:
        ob.cycle(sim, st)
    disp = np.mod(ob.pos[:, 0] - x.ravel() + 0.5, 1.0) - 0.5
    amp = np.abs(np.fft.rfft(disp.reshape(N, N, N), axis=-1)[..., 1]).mean() * 2 / N
    assert abs(amp / A - st.a / a0) < 0.05 * st.a / a0
    ob.close()
