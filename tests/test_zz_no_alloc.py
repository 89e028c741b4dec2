"""The allocation-free power spectrum (gev_extract_power_spectrum_noalloc)
against gev_extract_power_spectrum, through the C ABI: two calls in a row,
the second on a smaller box with fewer bins, each bit-identical to the
allocating entry point on the same field; sizes beyond the static scratch
are refused before any device work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _spectra(gev, N, numbins, seed, form):
    rng = np.random.default_rng(seed)
    H = N // 2 + 1
    F = rng.standard_normal((N, N, H)) + 1j * rng.standard_normal((N, N, H))
    with gev.Box(N, 8, form=form) as b:
        b.upload_field(gev.SCALARFT, F)
        want = b.extractPowerSpectrum(gev.SCALARFT, numbins)
        got = b.extractPowerSpectrum_noalloc(gev.SCALARFT, numbins)
        again = b.extractPowerSpectrum_noalloc(gev.SCALARFT, numbins)
    return want, got, again


@pytest.mark.parametrize("form", [0, 1])
def test_noalloc_spectrum_equals_the_allocating_one(gev, form):
    for N, numbins, seed in ((64, 32, 1), (32, 16, 2)):  # the second call is smaller
        want, got, again = _spectra(gev, N, numbins, seed, form)
        for w, g, a in zip(want, got, again):
            assert np.array_equal(w, g) and np.array_equal(g, a), (N, numbins)
        assert want[4].sum() > 0


def test_noalloc_spectrum_refuses_sizes_beyond_its_scratch(gev):
    with gev.Box(16, 8) as b:
        out = [np.zeros(300) for _ in range(4)] + [np.zeros(300, dtype=np.int32)]
        rc = gev.lib().gev_extract_power_spectrum_noalloc(b._c, gev.SCALARFT, 300, 1, 1,
                                                          *[o.ctypes.data for o in out])
        assert rc == -1 and b"256" in gev.lib().gev_last_error(b._c)
