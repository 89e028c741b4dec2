"""Pins the oracle's restatement of the basic IC generator (ic_basic.hpp) and
of its PRNG (prng_engine.hpp) against golden vectors, closed forms and
independent computations.  CPU only."""
import json
import math
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
S = 65536  # the IC streams' row stride, ic_basic.hpp:31-33


# ------------------------------------------------------------------ PRNG

def test_prng_matches_reference_golden_vectors():
    """Outputs of sitmo::prng_engine written by the reference header itself
    (seeds, discards within and across blocks, huge discards)."""
    g = json.load(open(os.path.join(GOLD, "sitmo_prng.json")))
    assert len(g["cases"]) >= 20
    for c in g["cases"]:
        pos, got = 0, []
        for st in c["steps"]:
            n = int(st[1:])
            if st[0] == "d":
                got += [O.prng_at(c["seed"], pos + i) for i in range(n)]
            pos += n
        assert got == c["values"], (c["seed"], c["steps"])


def test_prng_block_halves_order():
    """Output 2j is the low half and 2j+1 the high half of word j of a block."""
    for seed, blk in ((42, 0), (7, 12345), (0xFFFFFFFF, 2 ** 40)):
        w = O.prng_block(seed, blk)
        for j in range(4):
            assert O.prng_at(seed, 8 * blk + 2 * j) == int(w[j]) & 0xFFFFFFFF
            assert O.prng_at(seed, 8 * blk + 2 * j + 1) == int(w[j]) >> 32


def test_prng_uniformity():
    v = np.array([O.prng_at(99, i) for i in range(20000)], dtype=np.float64) / 2.0 ** 32
    assert abs(v.mean() - 0.5) < 0.01
    assert abs(v.var() - 1.0 / 12) < 0.005


# ------------------------------------------------------------------ spline

def test_spline_natural_matches_scipy():
    from scipy.interpolate import CubicSpline
    rng = np.random.default_rng(3)
    x = np.sort(rng.uniform(0, 10, 40))
    y = np.sin(x) + 0.1 * x ** 2
    c = O.spline_init(x, y)
    assert c[0] == 0.0 and c[-1] == 0.0
    cs = CubicSpline(x, y, bc_type="natural")
    xq = np.concatenate([rng.uniform(x[0], x[-1], 300), x])
    got = O.spline_eval(x, y, c, xq)
    assert np.max(np.abs(got - cs(xq))) < 1e-12 * np.max(np.abs(y))
    # second derivative at the knots = 2 c
    assert np.max(np.abs(2 * c - cs(x, 2))) < 1e-9


def test_spline_reproduces_linear_data_and_rejects_outside():
    x = np.array([0.5, 1.0, 2.5, 4.0, 7.0])
    y = 3.0 * x - 2.0
    c = O.spline_init(x, y)
    assert np.max(np.abs(c)) < 1e-14
    xq = np.linspace(0.5, 7.0, 50)
    assert np.max(np.abs(O.spline_eval(x, y, c, xq) - (3.0 * xq - 2.0))) < 1e-13
    assert np.isnan(O.spline_eval(x, y, c, [0.4999])[0])
    assert np.isnan(O.spline_eval(x, y, c, [7.0001])[0])


# ------------------------------------------------------------------ CIC kernel

def test_standard_kernel_and_its_transform():
    N = 16
    k = O.generate_cic_kernel(N)
    assert k[0, 0, 0] == 6 * N * N
    for (z, y, x) in ((0, 0, 1), (0, 0, N - 1), (0, 1, 0), (0, N - 1, 0), (1, 0, 0), (N - 1, 0, 0)):
        assert k[z, y, x] == -N * N
    assert np.count_nonzero(k) == 7
    # its transform is N^2 * sum_i (2 - 2 cos(2 pi k_i / N))
    ft = np.fft.rfftn(k)
    kz, ky, kx = np.meshgrid(np.arange(N), np.arange(N), np.arange(N // 2 + 1), indexing="ij")
    want = N * N * ((2 - 2 * np.cos(2 * np.pi * kx / N)) + (2 - 2 * np.cos(2 * np.pi * ky / N))
                    + (2 - 2 * np.cos(2 * np.pi * kz / N)))
    assert np.max(np.abs(ft - want)) < 1e-9 * N * N


def test_sc0_template_reproduces_standard_kernel_exactly():
    """A template whose particles sit on lattice points has the standard
    kernel as its CIC kernel (each direction: weight 1 at the point, q1 = 2)."""
    tpl = O.load_template(os.path.join(GOLD, "sc0_crystal.dat"))
    for N in (16, 32):
        assert np.array_equal(O.generate_cic_kernel(N, tpl, N // 4), O.generate_cic_kernel(N))


@pytest.mark.parametrize("name,ratio", [("sc1", 4), ("bcc", 4), ("fcc", 2)])
def test_template_kernels_sum_to_zero_and_are_local(name, ratio):
    tpl = O.load_template(os.path.join(GOLD, name + "_crystal.dat"))
    N = 16
    k = O.generate_cic_kernel(N, tpl, N // ratio)
    assert np.abs(k).max() > N * N
    assert abs(k.sum()) < 1e-9 * np.abs(k).max()
    nz = np.argwhere(k != 0)
    for c in nz:
        assert all(min(v, N - v) <= 1 for v in c), c


def test_sc1_kernel_is_reflection_symmetric():
    tpl = O.load_template(os.path.join(GOLD, "sc1_crystal.dat"))
    N = 16
    k = O.generate_cic_kernel(N, tpl, N // 4)
    for ax in range(3):
        r = np.roll(np.flip(k, axis=ax), 1, axis=ax)
        assert np.allclose(r, k, rtol=1e-12, atol=1e-9)


# ------------------------------------------------------------------ realisation

def drawn(kx, ky, kz, kmax, ksphere):
    if kx >= kmax or ky >= kmax or kz >= kmax:
        return False
    return not (ksphere and kx * kx + ky * ky + kz * kz >= kmax * kmax)


def line_layout(N):
    """For every (ky, kz) row of the k grid: the stream start of its kx >= kx0
    modes and, for kz > N/2, the stream start of its kx = 0 mode.  An
    independent statement of the layout of ic_basic.hpp:1089-1379."""
    H = N // 2
    out = {}
    for kz in range(N):
        for ky in range(N):
            yhi, zhi = ky > H, kz > H
            if not yhi and not zhi:
                out[ky, kz] = ((kz * S + ky) * S, 1 if ky == 0 and kz == 0 else 0, None)
            elif yhi and not zhi:
                out[ky, kz] = (((S + kz) * S + (N - ky)) * S, 0, None)
            elif not yhi:
                out[ky, kz] = (((2 * S + (N - kz)) * S + ky) * S, 1, ((N - kz) * S + ky) * S)
            else:
                out[ky, kz] = (((3 * S + (N - kz)) * S + (N - ky)) * S, 1,
                               ((2 * S + (N - kz)) * S + (N - ky)) * S)
    return out


def expected_unit_realisation(N, seed, ksphere):
    """Realisation with K = 1, P = 1, no deconvolution: e^{+-2 pi i r2}
    sqrt(-2 log r1), draws walked per row with the zero-redraw rule."""
    H, kmax = N // 2, N // 2 - 1
    out = np.zeros((N, N, H + 1), dtype=np.complex128)

    def walk(p):
        while True:
            u1 = O.prng_at(seed, p)
            p += 1
            if u1 != 0:
                break
        u2 = O.prng_at(seed, p)
        p += 1
        r1 = np.float32(u1) / np.float32(4294967296.0)
        r2 = np.float32(u2) / np.float32(4294967296.0)
        return p, float(r1), float(r2)

    for (ky, kz), (base, kx0, base0) in line_layout(N).items():
        yq, zq = (ky if ky <= H else N - ky), (kz if kz <= H else N - kz)
        p = base
        for kx in range(kx0, H + 1):
            if drawn(kx, yq, zq, kmax, ksphere):
                p, r1, r2 = walk(p)
                out[kz, ky, kx] = complex(math.cos(2 * math.pi * r2), math.sin(2 * math.pi * r2)) * \
                    math.sqrt(-2.0 * float(np.float32(math.log(r1))))
        if base0 is not None and drawn(0, yq, zq, kmax, ksphere):
            _, r1, r2 = walk(base0)
            out[kz, ky, 0] = complex(math.cos(2 * math.pi * r2), -math.sin(2 * math.pi * r2)) * \
                math.sqrt(-2.0 * float(np.float32(math.log(r1))))
    return out


def unit_spline():
    x = np.array([0.0, 1e6])
    y = np.array([1.0, 1.0])
    return x, y, O.spline_init(x, y)


@pytest.mark.parametrize("N,seed,ksphere", [(16, 42, 1), (16, 7, 0), (12, 3, 1)])
def test_realisation_stream_layout(N, seed, ksphere):
    x, y, c = unit_spline()
    H = N // 2
    got = O.generate_displacement_field(N, np.zeros((N, N, H + 1)), 0.0, x, y, c, seed, ksphere,
                                        deconvolve_f=0, ignorekernel=1, cr_log=1)
    want = expected_unit_realisation(N, seed, ksphere)
    assert np.array_equal(got == 0, want == 0)
    nz = want != 0
    assert np.max(np.abs(got[nz] - want[nz]) / np.abs(want[nz])) < 1e-6


def test_realisation_zero_redraw_path():
    """Seed found by tests/golden/find_rejection_seed.c: on a 16^3 grid with
    the sphere cut, mode (kx, ky, kz) = (0, 3, 2) draws r1 = 0 and redraws."""
    g = json.load(open(os.path.join(GOLD, "prng_rejection.json")))
    N, seed = g["N"], g["seed"]
    x, y, c = unit_spline()
    H = N // 2
    got = O.generate_displacement_field(N, np.zeros((N, N, H + 1)), 0.0, x, y, c, seed, g["ksphere"],
                                        deconvolve_f=0, ignorekernel=1, cr_log=1)
    want = expected_unit_realisation(N, seed, g["ksphere"])
    kx, ky, kz = g["mode"]
    assert O.prng_at(seed, ((kz * S + ky) * S) + 2 * kx) == 0
    nz = want != 0
    assert np.max(np.abs(got[nz] - want[nz]) / np.abs(want[nz])) < 1e-6


def test_realisation_amplitudes_follow_the_spectrum():
    """With the kernel and the deconvolution: E|F|^2 = 2 P(k)^2 s^2 |1/K|^2."""
    N = 32
    H = N // 2
    kx = np.arange(H + 1)
    kk = np.fft.fftfreq(N, 1.0 / N)
    KZ, KY, KX = np.meshgrid(np.abs(kk), np.abs(kk), kx, indexing="ij")
    ker = np.fft.rfftn(O.generate_cic_kernel(N))
    x = np.linspace(0.0, 2 * np.pi * N, 200)
    y = 1.0 / (1.0 + x)
    c = O.spline_init(x, y)
    F = O.generate_displacement_field(N, ker, 0.0, x, y, c, 11, 1)
    k2 = (KX ** 2 + KY ** 2 + KZ ** 2).astype(np.float32)
    sinc = np.ones(N)
    sinc[1:] = np.sin(np.pi * np.arange(1, N) / N) * N / (np.pi * np.arange(1, N))
    s = sinc[KX.astype(int)] * sinc[KY.astype(int)] * sinc[KZ.astype(int)]
    pk = 1.0 / (1.0 + 2 * np.pi * np.sqrt(k2))
    live = F != 0
    ratio = np.abs(F[live]) ** 2 / (pk[live] ** 2 * s[live] ** 2 / np.abs(ker[live]) ** 2)
    assert live.sum() > 5000
    assert abs(ratio.mean() / 2.0 - 1.0) < 0.05
    # exactly the modes inside the sphere |k| < N/2 - 1 are live, k = 0 excluded
    kmax = H - 1
    assert np.array_equal(live, (k2 < kmax * kmax) & (k2 > 0))


def test_glibc_logf_and_correctly_rounded_log_differ_by_at_most_one_ulp():
    """The reference's log(r1) is the C library's logf; the oracle's cr_log
    switch (and the HIP path) uses the correctly rounded float log.  Both
    realisations agree to one float ulp of the amplitude factor."""
    N = 16
    x, y, c = unit_spline()
    H = N // 2
    z = np.zeros((N, N, H + 1))
    a = O.generate_displacement_field(N, z, 0.0, x, y, c, 42, 1, 0, 1, cr_log=0)
    b = O.generate_displacement_field(N, z, 0.0, x, y, c, 42, 1, 0, 1, cr_log=1)
    nz = a != 0
    assert np.max(np.abs(a[nz] - b[nz]) / np.abs(b[nz])) < 2.0 ** -23


# ------------------------------------------------------------------ particles

def test_particle_positions_and_ids():
    tpl = O.load_template(os.path.join(GOLD, "sc1_crystal.dat"))
    nt = 2
    pos, ids = O.initialize_particle_positions(tpl, nt)
    assert len(pos) == 64 * nt ** 3
    assert np.array_equal(ids, np.arange(len(pos)))
    i = 5 + 64 * (1 + nt * (0 + nt * 1))  # template particle 5 of tile (x, y, z) = (1, 0, 1)
    assert np.array_equal(pos[i], (np.array([1.0, 0.0, 1.0]) + tpl[5].astype(np.float64)) / nt)
    # a tiled sc1 template is the regular lattice at cell centres
    g = np.sort(np.unique(pos[:, 0]))
    assert np.allclose(g, (np.arange(4 * nt) + 0.5) / (4 * nt), atol=1e-15)


def test_displacement_and_velocity_of_a_linear_field():
    """The CIC gradient of a field linear in the lattice index is exact: a
    particle away from the periodic seam moves by coeff * N * (gx, gy, gz)."""
    N = 16
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    gx, gy, gz = 1e-5, -2e-5, 3e-5
    f = (gx * x + gy * y + gz * z).astype(np.float64)
    rng = np.random.default_rng(1)
    pos = rng.uniform(1.0 / N, (N - 2.0) / N, (200, 3))
    p0 = pos.copy()
    md = O.displace_pcls_ic(N, pos, 0.7, f)
    d = pos - p0
    want = 0.7 * N * np.array([gx, gy, gz])
    assert np.max(np.abs(d - want)) < 1e-14
    assert abs(md - np.linalg.norm(want)) < 1e-15
    q, vmax = O.initialize_q_ic(N, p0, 2.0, f)
    assert np.max(np.abs(q + 2.0 * N * np.array([gx, gy, gz]))) < 1e-14
    assert abs(vmax - 2.0 * N * np.linalg.norm([gx, gy, gz])) < 1e-14


def test_displacement_wraps_into_the_box():
    N = 8
    f = np.zeros((N, N, N))
    f[:, :, 0] = 1e-3  # gradient at the seam
    pos = np.array([[0.99, 0.5, 0.5], [0.001, 0.5, 0.5]])
    O.displace_pcls_ic(N, pos, 1.0, f)
    assert np.all((pos >= 0.0) & (pos < 1.0))


# ------------------------------------------------------------------ tables and the whole generator

def ic_settings(N, seed=42, cr_log=1, boxsize=320.0):
    return O.IC(seed, 1, 1, N // 4, cr_log, 100.0, 2.1e-9, 0.965, 0.05, boxsize)


def test_transfer_loader_columns_and_units():
    k, d, t = O.load_transfer_functions(os.path.join(GOLD, "class_tk.dat"), "cdm", 320.0, 0.67556)
    assert len(k) == 135
    assert abs(k[0] - 1.045004358465e-05 * 320.0) < 1e-18
    assert d[0] == -1.163145528648e+00
    assert abs(t[0] - 1.505242433977e-08 * 320.0 / 0.67556) < 1e-20


def test_ic_tables_closed_forms():
    N = 16
    ic = ic_settings(N)
    cosmo = O.make_cosmo()
    sim = O.Sim(N, 1, 0, 48.0, 0.04, 30.0, 1.5 * 320.0 ** 2 / 2997.92458 ** 2, cosmo)
    tk = os.path.join(GOLD, "class_tk.dat")
    k, dt, tt = O.load_transfer_functions(tk, "tot", 320.0, cosmo.h)
    _, dc, tc = O.load_transfer_functions(tk, "cdm", 320.0, cosmo.h)
    yp, yd, yv = O.ic_tables(sim, ic, k, dt, tt, dc, tc)
    a = 1.0 / 101.0
    w = math.pi * np.sqrt(ic.A_s * (k * cosmo.h / 320.0 / ic.k_pivot) ** (ic.n_s - 1.0) / k) / k
    assert np.allclose(yv, -a * tc * w, rtol=1e-13, atol=0)
    assert np.allclose(yd, -3.0 * yp / k ** 2 - dc * w, rtol=1e-12, atol=0)
    # phi = (1.5 (H^2 - H0^2 a^2 Omega_L) d_tot + rescale t_tot / k^2) w, where rescale is
    # 3 H^3 (1 + H^2/2 d(1/H^2)/d ln a) by a five-point difference (ic_basic.hpp:1720); in
    # matter domination 1/H^2 grows like a, so rescale = 4.5 H^3 (z = 100: Lambda is 1e-6)
    Ha = O.hconf(a, sim.fourpiG, cosmo)
    H1 = O.hconf(1.0, sim.fourpiG, cosmo)
    A = 1.5 * (Ha * Ha - H1 * H1 * a * a * cosmo.Omega_Lambda)
    rescale = (yp / w - A * dt) * k * k / tt
    sel = np.abs(tt / k ** 2) > 1e-3 * np.abs(A * dt / Ha ** 3)
    assert sel.sum() > 20
    assert np.max(np.abs(rescale[sel] / (4.5 * Ha ** 3) - 1.0)) < 1e-4


@pytest.mark.parametrize("gr", [1, 0])
def test_generate_ic_basic_runs_and_is_consistent(gr):
    """The whole transfer-function path at 16^3: particle count, mass, the
    displacement bounded by the lattice, phi of primordial amplitude, and the
    same bits on a second run."""
    N = 16
    cosmo = O.make_cosmo()
    sim = O.Sim(N, gr, 0, 48.0, 0.04, 30.0, 1.5 * 320.0 ** 2 / 2997.92458 ** 2, cosmo)
    ic = ic_settings(N, boxsize=320.0)
    tpl = O.load_template(os.path.join(GOLD, "sc1_crystal.dat"))
    tk = os.path.join(GOLD, "class_tk.dat")
    k, dt, tt = O.load_transfer_functions(tk, "tot", 320.0, cosmo.h)
    _, dc, tc = O.load_transfer_functions(tk, "cdm", 320.0, cosmo.h)
    res = []
    for _ in range(2):
        b = O.OracleBox(N, N ** 3)
        mv, md = b.generate_ic(sim, ic, tpl, k, dt, tt, dc, tc)
        assert abs(b.box.pmass * N ** 3 - (cosmo.Omega_cdm + cosmo.Omega_b)) < 1e-15
        assert 0.0 < md * N < 1.0          # Zel'dovich displacements of a fraction of a cell at z = 100
        assert 0.0 < mv < 1e-2
        assert 1e-7 < np.abs(b.phi).max() < 1e-3
        assert np.all((b.pos >= 0) & (b.pos < 1))
        res.append((b.pos.copy(), b.q.copy(), b.phi.copy(), b.chi.copy(), b.Bi.copy(), mv, md))
        b.close()
    for x, y in zip(res[0], res[1]):
        assert np.array_equal(x, y)
