"""Pins the oracle's power-spectrum estimator (tools.hpp:53-212) and the
library's spectrum writer (tools.hpp:268-346).  CPU only."""
import numpy as np
import pytest

import oracle as O


def grid_tables(N, ktype, deconvolve):
    n = np.arange(N)
    if ktype == 0:
        tk = (2.0 * N * np.sin(np.pi * n / N)) ** 2
    else:
        tk = (2.0 * np.pi * np.minimum(n, N - n)) ** 2
    sinc = np.ones(N)
    if deconvolve:
        m = np.minimum(n, N - n)[1:].astype(np.float64)  # (float) i is exact; the expression is double
        sinc[1:] = np.sin(np.pi * m / N) * N / (np.pi * m)
    return tk, sinc


def numpy_spectrum(F, numbins, deconvolve, ktype, pfun):
    """Independent vectorised binning of the same estimator."""
    N = F.shape[-2]
    H = N // 2
    tk, sinc = grid_tables(N, ktype, deconvolve)
    kz, ky, kx = np.meshgrid(np.arange(N), np.arange(N), np.arange(H + 1), indexing="ij")
    k2 = tk[kx] + tk[ky] + tk[kz]
    s = (sinc[kx] * sinc[ky] * sinc[kz]) ** 2
    w = np.where((kx == 0) | (kx == H), 1, 2)
    live = ~((kx == 0) & (ky == 0) & (kz == 0))
    b = np.floor(numbins * np.sqrt(k2 / (3.0 * tk[H]))).astype(int)
    live &= b < numbins
    p = pfun(F)
    occ = np.bincount(b[live], weights=w[live], minlength=numbins)
    ksum = np.bincount(b[live], weights=(w * np.sqrt(k2))[live], minlength=numbins)
    k2sum = np.bincount(b[live], weights=(w * k2)[live], minlength=numbins)
    psum = np.bincount(b[live], weights=(w * p * k2 * np.sqrt(k2) / s)[live], minlength=numbins)
    p2sum = np.bincount(b[live], weights=(w * p * p * k2 ** 3 / s ** 2)[live], minlength=numbins)
    with np.errstate(invalid="ignore", divide="ignore"):
        kbin = np.where(occ > 0, ksum / occ, 0)
        power = np.where(occ > 0, psum / occ, 0)
        # scatters as weighted standard deviations, two-pass (well conditioned)
        kd = np.sqrt(k2)[live] - kbin[b[live]]
        pd = (p * k2 * np.sqrt(k2) / s)[live] - power[b[live]]
        kstd = np.sqrt(np.bincount(b[live], weights=w[live] * kd * kd, minlength=numbins) / occ)
        pstd = np.sqrt(np.bincount(b[live], weights=w[live] * pd * pd, minlength=numbins) / occ)
    kstd = np.where(occ > 0, kstd, 0)
    pstd = np.where(occ > 0, pstd, 0)
    return kbin, power, occ.astype(np.int32), kstd, pstd


@pytest.mark.parametrize("N,numbins,deconvolve,ktype", [(16, 16, 1, 1), (24, 40, 0, 1), (16, 8, 1, 0), (32, 64, 1, 1)])
def test_scalar_spectrum_matches_numpy_binning(N, numbins, deconvolve, ktype):
    rng = np.random.default_rng(N + numbins)
    F = np.fft.rfftn(rng.standard_normal((N, N, N)))
    kb, pw, ks, ps, oc = O.extract_cross_spectrum(F, numbins=numbins, deconvolve=deconvolve, ktype=ktype)
    kb2, pw2, oc2, ks2, ps2 = numpy_spectrum(F, numbins, deconvolve, ktype, lambda f: np.abs(f) ** 2)
    assert np.array_equal(oc, oc2)
    assert np.allclose(kb, kb2, rtol=1e-13, atol=0)
    assert np.allclose(pw, pw2, rtol=1e-11, atol=0)
    # the one-pass scatter of tools.hpp:188-192 is the weighted standard
    # deviation up to its cancellation error (~sqrt(eps) x the bin mean)
    assert np.all(np.abs(ks - ks2) <= 1e-6 * kb)
    assert np.all(np.abs(ps - ps2) <= 1e-6 * pw)
    assert np.all(ks >= 0) and np.all(ps >= 0)


def test_single_mode_closed_form():
    """One mode A at k0: power in its bin is w |A|^2 k0^3 / s / occupation."""
    N, numbins = 16, 16
    F = np.zeros((N, N, N // 2 + 1), dtype=np.complex128)
    F[1, 2, 3] = 3.0 - 4.0j  # (kz, ky, kx) = (1, 2, 3)
    kb, pw, ks, ps, oc = O.extract_cross_spectrum(F, numbins=numbins, deconvolve=0, ktype=1)
    k2 = (2 * np.pi) ** 2 * (9 + 4 + 1)
    b = int(np.floor(numbins * np.sqrt(k2 / (3 * (2 * np.pi * 8) ** 2))))
    assert pw[b] == pytest.approx(2 * 25.0 * k2 ** 1.5 / oc[b], rel=1e-14)
    assert np.count_nonzero(pw) == 1
    # occupation counts every mode of the half grid with its weight (2 for 0 < kx < N/2)
    assert oc.sum() == N ** 3 - 1 - np.count_nonzero(
        np.floor(numbins * np.sqrt(np.add.outer(np.add.outer(*(np.minimum(np.arange(N), N - np.arange(N)) ** 2,) * 2),
                                                  np.minimum(np.arange(N), N - np.arange(N)) ** 2) / (3 * 64))) >= numbins)


def test_tensor_and_vector_sums():
    """Symmetric tensor: off-diagonals count twice (the full 3x3 |S_ij|^2);
    a vector sums its components; comp1/comp2 pick one pair."""
    N = 16
    rng = np.random.default_rng(5)
    S = np.fft.rfftn(rng.standard_normal((6, N, N, N)), axes=(1, 2, 3))
    full = np.abs(S[0]) ** 2 + np.abs(S[3]) ** 2 + np.abs(S[5]) ** 2 + 2 * (np.abs(S[1]) ** 2 + np.abs(S[2]) ** 2 + np.abs(S[4]) ** 2)
    _, pw, _, _, oc = O.extract_cross_spectrum(S, numbins=12)
    _, pw2, oc2, _, _ = numpy_spectrum(S, 12, 1, 1, lambda f: full)
    assert np.array_equal(oc, oc2) and np.allclose(pw, pw2, rtol=1e-11)
    V = S[:3].copy()
    _, pv, _, _, _ = O.extract_cross_spectrum(V, numbins=12)
    _, pv2, _, _, _ = numpy_spectrum(V, 12, 1, 1, lambda f: (np.abs(f) ** 2).sum(axis=0))
    assert np.allclose(pv, pv2, rtol=1e-11)
    _, pc, _, _, _ = O.extract_cross_spectrum(S, S, numbins=12, comp1=1, comp2=4)
    _, pc2, _, _, _ = numpy_spectrum(S, 12, 1, 1, lambda f: (f[1] * np.conj(f[4])).real)
    assert np.allclose(pc, pc2, rtol=1e-10, atol=1e-9 * np.abs(pc2).max())


def test_writer_format_and_exact_redshift_interpolation(gev, tmp_path):
    nb = 6
    kbin = np.linspace(1.0, 6.0, nb)
    occ = np.array([3, 0, 5, 7, 0, 2], dtype=np.int32)
    p1 = np.array([10.0, 0.0, 30.0, 40.0, 0.0, 60.0])
    spec1 = (kbin, p1, 0.1 * kbin, 0.5 * p1, occ)
    f = str(tmp_path / "pk.dat")
    # before the target redshift: plain table
    gev.writePowerSpectrum(spec1, 2.0, 4.0, f, "power spectrum of delta", 1.0 / 2.2, z_target=1.0)
    lines = open(f).read().splitlines()
    assert lines[0] == "# power spectrum of delta"
    assert lines[1] == "# redshift z=1.200000"
    rows = np.loadtxt(f)
    assert rows.shape == (4, 5) and np.array_equal(rows[:, 4], occ[occ > 0])
    assert np.allclose(rows[:, 0], kbin[occ > 0] / 2.0) and np.allclose(rows[:, 1], p1[occ > 0] / 4.0, rtol=1e-6)
    assert np.allclose(rows[:, 3], 0.5 * p1[occ > 0] / 4.0 / np.sqrt(occ[occ > 0]), rtol=1e-6)
    # after crossing z_target = 1: linear interpolation in redshift to z = 1
    p2 = 2.0 * p1
    a2 = 1.0 / 1.9
    gev.writePowerSpectrum((kbin, p2, 0.1 * kbin, 0.5 * p2, occ), 2.0, 4.0, f, "power spectrum of delta", a2, 1.0)
    w = (1.2 - 1.0) / (1.0 + 1.2 - 1.0 / a2)
    rows2 = np.loadtxt(f)
    assert open(f).read().splitlines()[1] == "# redshift z=1.000000"
    want = (1 - w) * np.round(p1 / 4.0, 6) + w * p2 / 4.0
    assert np.allclose(rows2[:, 1], want[occ > 0], rtol=1e-6)
