"""Pins where the oracle puts the staggered quantities, with one particle on
an 8^3 grid and a phi (or B) that is linear in the lattice coordinates, so
that every average the reference takes over a plaquette or a link equals the
linear function at that plaquette's or link's centre:

- T_ij (i != j) lives on the plaquette x + (e_i + e_j)/2, linear only in the
  third axis, with the factor 1 + f * (the plaquette's mean phi)
  (gevolution.hpp:1262-1272, scatter 1277-1294);
- T0i lives on the link x + e_i/2, NGP along i and linear in the other two
  axes, with the factor 1 + phi_a + phi_b of the link's ends
  (gevolution.hpp:1129-1144);
- the drift's B term interpolates B_i from the four i-links around the
  particle's cell, bilinearly in the two other axes (gevolution.hpp:852-865).

Each test also shows that the same expectation with the plaquette or link
moved by one cell along any axis does not match, so a shifted placement in
the oracle (and so in the GPU kernels, which equal it bit for bit) fails.
CPU only."""
import itertools

import numpy as np

import oracle as O

N = 8


def linear(coeffs):
    """f(x, y, z) = c0 + c1 x + c2 y + c3 z in lattice units, and its grid
    array indexed [z, y, x]."""
    c0, c1, c2, c3 = coeffs

    def f(x, y, z):
        return c0 + c1 * x + c2 * y + c3 * z

    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    return f, np.ascontiguousarray(f(x, y, z).astype(np.float64))


def particle(cell, frac, qv):
    c = np.asarray(cell, dtype=float)
    pos = np.array([(c + np.asarray(frac)) / N])
    return pos, np.array([qv], dtype=float), np.array([0], dtype=np.int64)


CELL = (3, 4, 2)              # away from the box edge: no periodic wrap in play
FRAC = (0.3125, 0.625, 0.1875)  # exact in binary
PHI = (2e-3, 1.1e-4, -2.3e-4, 3.7e-4)


def tij_expected(cell, frac, q, a, m, phif, shift=(0, 0, 0)):
    """T_01, T_02, T_12 of one particle from the plaquette geometry; `shift`
    moves every plaquette centre by whole cells (the counter-checks)."""
    cx, cy, cz = cell
    ux, uy, uz = frac
    sx, sy, sz = shift
    q2 = float(np.dot(q, q))
    e = np.sqrt(q2 + a * a)
    f = 4.0 + a * a / (q2 + a * a)
    mass = m * N ** 3 / a
    out = {}
    # (component, first index pair, the linear axis, its lower weight)
    w01, w02, w12 = mass * q[0] * q[1] / e, mass * q[0] * q[2] / e, mass * q[1] * q[2] / e
    for t in (0, 1):
        wz, wy, wx = (uz if t else 1 - uz), (uy if t else 1 - uy), (ux if t else 1 - ux)
        out[(1, cz + t, cy, cx)] = w01 * wz * (1 + f * phif(cx + 0.5 + sx, cy + 0.5 + sy, cz + t + sz))
        out[(2, cz, cy + t, cx)] = w02 * wy * (1 + f * phif(cx + 0.5 + sx, cy + t + sy, cz + 0.5 + sz))
        out[(4, cz, cy, cx + t)] = w12 * wx * (1 + f * phif(cx + t + sx, cy + 0.5 + sy, cz + 0.5 + sz))
    return out


def test_tij_offdiagonals_live_on_plaquettes():
    a, m, qv = 0.7, 1e-3, (0.011, -0.007, 0.013)
    pos, q, ids = particle(CELL, FRAC, qv)
    phif, phi = linear(PHI)
    T = O.projection_Tij_project(N, pos, q, ids, m, a=a, phi=phi)
    want = tij_expected(CELL, FRAC, np.array(qv), a, m, phif)
    for comp in (1, 2, 4):
        nz = {k[1:]: v for k, v in want.items() if k[0] == comp}
        got = T[comp]
        assert np.count_nonzero(got) == 2
        for (z, y, x), v in nz.items():
            assert abs(got[z, y, x] - v) <= 1e-14 * abs(v), (comp, (x, y, z), got[z, y, x], v)
    # the plaquette moved by one cell along any axis is a different answer
    for sh in ((1, 0, 0), (0, 1, 0), (0, 0, 1), (-1, 0, 0), (0, -1, 0), (0, 0, -1)):
        moved = tij_expected(CELL, FRAC, np.array(qv), a, m, phif, shift=sh)
        worst = max(abs(T[k[0], k[1], k[2], k[3]] - v) / abs(v) for k, v in moved.items())
        assert worst > 1e-6, sh


def test_tij_diagonals_on_vertices():
    a, m, qv = 0.4, 1e-3, (0.02, 0.005, -0.015)
    pos, q, ids = particle(CELL, FRAC, qv)
    phif, phi = linear(PHI)
    T = O.projection_Tij_project(N, pos, q, ids, m, a=a, phi=phi)
    q = np.array(qv)
    q2 = float(q @ q)
    e, f, mass = np.sqrt(q2 + a * a), 4.0 + a * a / (q2 + a * a), m * N ** 3 / a
    for comp, i in ((0, 0), (3, 1), (5, 2)):
        for dx, dy, dz in itertools.product((0, 1), repeat=3):
            W = np.prod([FRAC[k] if d else 1 - FRAC[k] for k, d in enumerate((dx, dy, dz))])
            x, y, z = CELL[0] + dx, CELL[1] + dy, CELL[2] + dz
            v = mass * q[i] * q[i] / e * W * (1 + f * phif(x, y, z))
            assert abs(T[comp, z, y, x] - v) <= 1e-14 * abs(v)
        assert np.count_nonzero(T[comp]) == 8


def t0i_expected(cell, frac, q, m, phif, shift=(0, 0, 0)):
    cx, cy, cz = cell
    ux, uy, uz = frac
    sx, sy, sz = shift
    mass = m * N ** 3  # coeff * m / dx^3, no 1/a (gevolution.hpp:1064-1065)
    out = {}
    for s, t in itertools.product((0, 1), repeat=2):
        # component 0: the x-link from (cx, y, z) to (cx + 1, y, z)
        y, z = cy + s, cz + t
        out[(0, z, y, cx)] = mass * q[0] * (uy if s else 1 - uy) * (uz if t else 1 - uz) * \
            (1 + 2 * phif(cx + 0.5 + sx, y + sy, z + sz))
        x, z = cx + s, cz + t
        out[(1, z, cy, x)] = mass * q[1] * (ux if s else 1 - ux) * (uz if t else 1 - uz) * \
            (1 + 2 * phif(x + sx, cy + 0.5 + sy, z + sz))
        x, y = cx + s, cy + t
        out[(2, cz, y, x)] = mass * q[2] * (ux if s else 1 - ux) * (uy if t else 1 - uy) * \
            (1 + 2 * phif(x + sx, y + sy, cz + 0.5 + sz))
    return out


def test_t0i_lives_on_links():
    m, qv = 1e-3, (0.011, -0.007, 0.013)
    pos, q, ids = particle(CELL, FRAC, qv)
    phif, phi = linear(PHI)
    T = O.projection_T0i_project(N, pos, q, ids, m, phi=phi)
    want = t0i_expected(CELL, FRAC, np.array(qv), m, phif)
    for comp in range(3):
        assert np.count_nonzero(T[comp]) == 4
    for k, v in want.items():
        assert abs(T[k] - v) <= 1e-14 * abs(v), (k, T[k], v)
    for sh in ((1, 0, 0), (0, 1, 0), (0, 0, 1), (-1, 0, 0), (0, -1, 0), (0, 0, -1)):
        moved = t0i_expected(CELL, FRAC, np.array(qv), m, phif, shift=sh)
        assert max(abs(T[k] - v) / abs(v) for k, v in moved.items()) > 1e-6, sh


def drift_b(frac, bfs, shift=(0, 0, 0)):
    """b_i of update_pos: B_i interpolated at the particle's position along
    the two axes other than i, at the i-link of the particle's cell."""
    cx, cy, cz = CELL
    rx, ry, rz = frac
    sx, sy, sz = shift
    return np.array([bfs[0](cx + sx, cy + ry + sy, cz + rz + sz),
                     bfs[1](cx + rx + sx, cy + sy, cz + rz + sz),
                     bfs[2](cx + rx + sx, cy + ry + sy, cz + sz)])


def test_drift_b_term_reads_the_four_links():
    """q = 0, phi = chi = 0: the drift is dtau * b / (a^2 N) with b from
    the links; each B_i is linear along every axis, so the bilinear link
    interpolation equals B_i at the particle's position on its own link."""
    coeffs = [(1e-4, 2e-5, -3e-5, 5e-5), (-2e-4, 4e-5, 1e-5, -2e-5), (3e-4, -1e-5, 6e-5, 2e-5)]
    fs, arrs = zip(*[linear(c) for c in coeffs])
    Bi = np.ascontiguousarray(np.stack(arrs))
    pos, q, _ = particle(CELL, FRAC, (0.0, 0.0, 0.0))
    zero = np.zeros((N, N, N))
    a, dtau = 0.5, 0.01
    params = (a, a * a * N)
    new = O.move_particles(N, pos, q, O.POS_GR, dtau, zero, zero, Bi, 3, params)
    got = (new[0] - pos[0])
    want = dtau * drift_b(FRAC, fs) / params[1]
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=0)
    for sh in ((1, 0, 0), (0, 1, 0), (0, 0, 1), (-1, 0, 0), (0, -1, 0), (0, 0, -1)):
        moved = dtau * drift_b(FRAC, fs, shift=sh) / params[1]
        assert np.max(np.abs(got - moved) / np.abs(moved)) > 1e-3, sh
    # without the B field (nfield 2) there is no drift at q = 0
    still = O.move_particles(N, pos, q, O.POS_GR, dtau, zero, zero, Bi, 2, params)
    assert np.array_equal(still, pos)
