"""Host model of the device kernels' index arithmetic: every global index a
kernel forms from its tile, staging region, column or row must lie inside
the buffer it reads or writes, for every grid size with a compiled FFT plan
(8 ... 1024).  The model restates the kernels' integer expressions axis by
axis (the staged regions are products of per-axis ranges), with the tile
shapes read from the kernel sources' defaults, so it follows them when a
default changes.  CPU only.

Kernels covered (gevolution-1.2_amd/csrc):
  particles.hip  k_update_tiled   field staging of phi, chi and B0..B2,
                                  row extents start[g0], start[g0 + nx]
  deposit.hip    k_deposit_tiled  target neighbours, source region slots,
                                  the row fast path's start[g + SX]
  fft.hip        k_r2c_x / k_c2r_x rows, k_c2c_cols columns (y and z)
  fftsolve.hip   k_zsolve columns of every kind
  spectrum.hip   k_spectrum_rows row ranges and the per-row LDS arrays
"""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "gevolution-1.2_amd", "csrc")
SIZES = (8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024)


def _define(fname, name):
    with open(os.path.join(CSRC, fname)) as f:
        m = re.search(r"#define\s+%s\s+(\d+)" % name, f.read())
    assert m, (fname, name)
    return int(m.group(1))


def _sizes_in_header():
    with open(os.path.join(CSRC, "gev_internal.h")) as f:
        text = f.read()
    m = re.search(r"#define GEV_FFT_SIZES\(X\)\s*\\\s*\n(.*)", text)
    return tuple(int(v) for v in re.findall(r"X\((\d+)\)", m.group(1)))


def test_model_covers_every_compiled_size():
    assert _sizes_in_header() == SIZES


def wrapi(i, n):
    """particles.hip / deposit.hip wrapi: one periodic step."""
    return i + n if i < 0 else (i - n if i >= n else i)


# ---------------------------------------------------------------- kick / drift

KTX = _define("particles.hip", "GEV_KICK_TX")
KTY = _define("particles.hip", "GEV_KICK_TY")
KTZ = _define("particles.hip", "GEV_KICK_TZ")


def kick_axis_coords(N, T, lo_here, clamp=True):
    """Coordinates one axis of a staged region reaches over every tile:
    slot r in [0, T + 1 + lo_here) maps to x0 - lo_here + min(r, n + lo_here)
    (n = the tile's extent), then one periodic step."""
    out = []
    for x0 in range(0, N, T):
        n = min(T, N - x0)
        for r in range(T + 1 + lo_here):
            rr = min(r, n + lo_here) if clamp else r
            out.append(wrapi(x0 - lo_here + rr, N))
    return out


@pytest.mark.parametrize("N", SIZES)
@pytest.mark.parametrize("LO", (-1, 0, 1, 2))
def test_kick_staging_in_bounds(N, LO):
    for axis, T in enumerate((KTX, KTY, KTZ)):
        c = kick_axis_coords(N, T, 1 if LO == axis else 0)
        assert min(c) >= 0 and max(c) < N, (N, LO, axis, min(c), max(c))


@pytest.mark.parametrize("N", SIZES)
def test_kick_staging_covers_what_the_particles_read(N):
    """Every cell a particle of the tile reads (offsets -1 on axis LO, 0 and
    +1) is a staged slot that holds that cell."""
    for LO in (-1, 0, 1, 2):
        for axis, T in enumerate((KTX, KTY, KTZ)):
            lo = 1 if LO == axis else 0
            for x0 in range(0, N, T):
                n = min(T, N - x0)
                staged = [wrapi(x0 - lo + min(r, n + lo), N) for r in range(T + 1 + lo)]
                for c in range(x0, x0 + n):
                    for d in ((-1, 0, 1) if lo else (0, 1)):
                        slot = c - x0 + lo + d
                        assert staged[slot] == (c + d) % N, (N, LO, axis, x0, c, d)


def test_model_detects_the_unclamped_overshoot():
    """Without the clamp, tiles wider than the box (N < KTX) stage past N."""
    bad = [N for N in SIZES if max(kick_axis_coords(N, KTX, 0, clamp=False)) >= N]
    assert bad == [8, 12, 16]


@pytest.mark.parametrize("N", SIZES)
def test_kick_row_extents_in_bounds(N):
    """start has N^3 + 1 entries; row r reads start[g0] and start[g0 + nx]."""
    n3 = N ** 3
    worst = 0
    for x0 in range(0, N, KTX):
        nx = min(KTX, N - x0)
        worst = max(worst, x0 + nx)  # g0 + nx at y = z = N - 1
    assert (N - 1) * N + (N - 1) * N * N + worst <= n3


# ---------------------------------------------------------------- deposit

TX = _define("deposit.hip", "GEV_DEP_TX")
TY = _define("deposit.hip", "GEV_DEP_TY")
TZ = _define("deposit.hip", "GEV_DEP_TZ")


def order_slot(p, lo0, n):
    return p if lo0 > 0 else (0 if p == n - 1 else p + 1)


@pytest.mark.parametrize("N", SIZES)
def test_deposit_source_region_in_bounds(N):
    for T in (TX, TY, TZ):
        for x0 in range(0, N, T):
            n = min(T, N - x0)
            ns = n + 1
            for p in range(ns):
                g = wrapi(x0 - 1 + order_slot(p, x0, ns), N)
                assert 0 <= g < N, (N, T, x0, p, g)
            # target neighbours xp = x == N - 1 ? 0 : x + 1
            for x in range(x0, x0 + n):
                xp = 0 if x == N - 1 else x + 1
                assert 0 <= xp < N


@pytest.mark.parametrize("N", SIZES)
def test_deposit_row_fast_path_in_bounds(N):
    """Full interior tiles (N a multiple of the tile, x0, y0, z0 > 0) read
    start[g .. g + SX] with g = L(x0 - 1, y, z): at most N^3."""
    if N % TX or N % TY or N % TZ or N < 2 * max(TX, TY, TZ):
        return
    SX = TX + 1
    n3 = N ** 3
    for x0 in range(TX, N, TX):
        for y in (TY - 1, N - 1):
            for z in (TZ - 1, N - 1):
                g = (x0 - 1) + N * (y + N * z)
                assert g + SX <= n3


# ---------------------------------------------------------------- FFT passes

def kpitch(N):
    return ((N // 2 + 1) + 7) // 8 * 8


def x_lines(N):
    m, l = N // 2, 1
    while l * 2 * m <= 2048:
        l *= 2
    return l


def col_tile(N, elem):
    w = 128 // (2 * elem)
    while w > 1 and w * N * 2 * elem > 65536:
        w //= 2
    return w


@pytest.mark.parametrize("N", SIZES)
@pytest.mark.parametrize("elem", (8, 4))
def test_fft_passes_in_bounds(N, elem):
    P, H = kpitch(N), N // 2 + 1
    nk = P * N * N
    # x passes: rows guarded by row < nrows; a row reads n < N/2 pairs
    # (r2c) or k, M - k <= N/2 < P (c2r)
    assert H <= P
    # column passes: blockIdx.x * W + w < H is checked, e < N, blockIdx.y < N
    W = col_tile(N, elem)
    ntx = (H + W - 1) // W
    for es, os_ in ((P, P * N), (P * N, P)):  # y pass, z pass
        last = (N - 1) * os_ + (H - 1) + (N - 1) * es
        assert last < nk
        assert (ntx - 1) * W < H


def zs_cols(N, tensor, f32):
    if N >= 768:
        return 2
    if f32:
        return 8 if tensor else 4
    return 4


@pytest.mark.parametrize("N", SIZES)
@pytest.mark.parametrize("f32", (False, True))
@pytest.mark.parametrize("nin", (1, 3, 6))
def test_zsolve_columns_in_bounds(N, f32, nin):
    P, H = kpitch(N), N // 2 + 1
    nk = P * N * N
    W = zs_cols(N, nin == 6, f32)
    # the last column a workgroup of the last tile row reads (kx < H checked)
    last = (nin - 1) * nk + (N - 1) * P + (H - 1) + (N - 1) * P * N
    assert last < nin * nk
    ntx = (H + W - 1) // W
    assert ntx * N < 2 ** 31


# ---------------------------------------------------------------- spectrum

@pytest.mark.parametrize("N", SIZES)
def test_spectrum_rows_in_bounds(N):
    rows = N * N
    rpb = (rows + 1023) // 1024
    nblk = (rows + rpb - 1) // rpb
    covered = sum(min(b * rpb + rpb, rows) - b * rpb for b in range(nblk))
    assert covered == rows
    maxh = 257 if N // 2 + 1 <= 257 else 513
    assert N // 2 + 1 <= maxh  # per-row LDS arrays hold kx = 0 .. N/2
    P = kpitch(N)
    assert (rows - 1) * P + N // 2 < P * N * N
