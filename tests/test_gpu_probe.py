"""The bandwidth probe (hand-written streaming kernels) through the C ABI:
every kind returns a rate below the 8 TB/s peak and above 1 TB/s, and a
copy moves twice the bytes of a read in about twice the time."""
import pytest

pytestmark = pytest.mark.gpu


def test_bandwidth_probe_rates(gev):
    rates = {}
    for name, kind in (("copy", gev.PROBE_COPY), ("copy_nt", gev.PROBE_COPY_NT), ("read", gev.PROBE_READ),
                       ("write", gev.PROBE_WRITE)):
        gbs, ms = gev.bandwidth_probe(1 << 30, kind, 5)
        assert 1000.0 < gbs < 8000.0 and ms > 0, (name, gbs, ms)
        rates[name] = (gbs, ms)
    assert 1.2 < rates["copy"][1] / rates["read"][1] < 3.5
    print({k: round(v[0]) for k, v in rates.items()}, flush=True)
