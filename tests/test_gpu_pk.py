"""GPU parity of the power-spectrum estimator (tools.hpp:53-212) against the
oracle, and the GR delta spectrum sequence of output.hpp:1990-2028.  The
device sums per bin in a fixed order different from the oracle's sequential
one, so the bin means (k, P) agree to 1e-12 relative.  The scatters are
sqrt(<x^2> - <x>^2): where a bin's modes share one value (true scatter 0)
any summation order leaves sqrt(rounding noise), about 1e-7 of the mean, so
they are compared with |difference| <= 1e-6 x the bin mean (the bar for P(k)).
Occupations are exact and repeated calls give identical bits."""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def close(a, b, rtol):
    a, b = np.asarray(a), np.asarray(b)
    return np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(b), 1e-300 + 1e-14 * np.abs(b).max()))


def assert_spectra_match(g, o, rtol_mean):
    """(kbin, power, kscatter, pscatter, occupation) of the device vs the oracle."""
    assert np.array_equal(g[4], o[4])
    assert close(g[0], o[0], rtol_mean)
    assert close(g[1], o[1], rtol_mean)
    assert np.all(np.abs(g[2] - o[2]) <= 1e-6 * np.abs(o[0])), "kscatter"
    assert np.all(np.abs(g[3] - o[3]) <= 1e-6 * np.abs(o[1])), "pscatter"


@pytest.mark.parametrize("N,numbins,deconvolve,ktype", [(32, 64, 1, 1), (48, 100, 0, 0), (16, 1024, 1, 1),
                                                        (64, 256, 1, 1)])
def test_scalar_spectrum(gev, N, numbins, deconvolve, ktype):
    rng = np.random.default_rng(N)
    F = np.fft.rfftn(rng.standard_normal((N, N, N)))
    with gev.Box(N, 8) as b:
        b.upload_field(gev.SCALARFT, F)
        g = b.extractPowerSpectrum(gev.SCALARFT, numbins, deconvolve, ktype)
        g2 = b.extractPowerSpectrum(gev.SCALARFT, numbins, deconvolve, ktype)
    o = O.extract_cross_spectrum(F, numbins=numbins, deconvolve=deconvolve, ktype=ktype)
    assert_spectra_match(g, o, 1e-12)
    for i in range(5):
        assert np.array_equal(g[i], g2[i])


def test_tensor_vector_and_cross_spectra(gev):
    N = 24
    rng = np.random.default_rng(3)
    S = np.fft.rfftn(rng.standard_normal((6, N, N, N)), axes=(1, 2, 3))
    B = np.fft.rfftn(rng.standard_normal((3, N, N, N)), axes=(1, 2, 3))
    with gev.Box(N, 8) as b:
        b.upload_field(gev.SIJFT, S)
        b.upload_field(gev.BIFT, B)
        gs = b.extractPowerSpectrum(gev.SIJFT, 40)
        gv = b.extractPowerSpectrum(gev.BIFT, 40)
        gc = b.extractCrossSpectrum(gev.SIJFT, gev.BIFT, 40, comp1=4, comp2=1)
        with pytest.raises(gev.GevError):
            b.extractCrossSpectrum(gev.SIJFT, gev.BIFT, 40)  # 6 vs 3 components
    os_ = O.extract_cross_spectrum(S, numbins=40)
    ov = O.extract_cross_spectrum(B, numbins=40)
    oc = O.extract_cross_spectrum(S, np.concatenate([B, B]), numbins=40, comp1=4, comp2=1)
    for g, o in ((gs, os_), (gv, ov), (gc, oc)):
        assert_spectra_match(g, o, 1e-11)


def test_delta_spectrum_of_generated_ics(gev):
    """output.hpp:1990-2028 on the IC state: projection_T00 with phi, FFT,
    extractPowerSpectrum(deconvolve, linear k), rescaled as writePowerSpectrum
    does for the delta spectrum; device and oracle ICs give the same P(k)."""
    from test_gpu_ic import ic_inputs
    N, boxsize = 32, 160.0
    cosmo_o, fpg, tabs, tpl, sim_o, ic_o = ic_inputs(N, boxsize, 1)
    ob = O.OracleBox(N, N ** 3)
    ob.generate_ic(sim_o, ic_o, tpl, *tabs)
    a = 1.0 / 101.0
    T00 = O.projection_T00_project(N, ob.pos, ob.q, ob.id, ob.box.pmass, a=a, phi=ob.phi)
    o = O.extract_cross_spectrum(O.fft_r2c(T00), numbins=32)
    sim_g = gev.SimParams(1, 0, 48.0, 0.04, 30.0, fpg, gev.make_cosmology())
    ic_g = gev.ICParams(ic_o.seed, 1, 1, N // 4, 100.0, 2.1e-9, 0.965, 0.05, boxsize)
    with gev.Box(N, N ** 3) as b:
        b.generateIC_basic(sim_g, ic_g, tpl, *tabs)
        b.projection_init(gev.SOURCE)
        b.projection_T00_project(a, True, 1.0)
        b.fft(gev.PLAN_SOURCE, gev.FFT_FORWARD)
        g = b.extractPowerSpectrum(gev.SCALARFT, 32)
    ob.close()
    assert np.array_equal(g[4], o[4])
    live = o[4] > 0
    assert close(g[1][live], o[1][live], 1e-8)
    rescalep = float(N ** 3) ** 2 * 2 * np.pi ** 2 * 0.31 ** 2
    assert np.all(g[1][1:][live[1:]] / rescalep > 0)
