"""The FP32 precision form (gev_box_desc.form = GEV_FORM_FP32) against the
plain FP64 form, which the oracle pins.

The FP32 form stores phi, chi, source, Sij, Bi and the k-space fields as
float and runs the FFT butterflies, the deposit's sums, the tensor stencil,
the kick and drift gathers and the tensor solves' field combinations in
float; the particles, the background, every factor involving a particle's
momentum, the scalar stencil, and the k-space solves' coefficients,
denominators and zero mode stay FP64.  Its results therefore differ from the
FP64 form by float rounding, and the bars are stated per quantity:
  * one FFT round trip and a forward transform: max error <= 2e-6 x max |ref|
    (about 17 ulp of float at the magnitudes involved, log2 N stages);
  * whole cycles from the same state: phi, chi within 1e-4 relative (max
    norm; phi(k=0), the mean of phi, within 1e-4 of max |phi|), Bi within
    1e-3 (a small field built from differences), momenta within 1e-5 of
    their change over the run, positions within 1e-9 absolute (box units);
  * the matter power spectrum (T00 with phi, FFT, extractPowerSpectrum) of
    generated LCDM initial conditions evolved in both forms: within 1e-4
    relative in every bin below k_Nyq/2;
  * run to run the FP32 form is bit-reproducible (the deposit is a gather in
    a fixed order, with no atomics), and its fused cycle equals its own call
    sequence bit for bit, as the FP64 form's does.
"""
import numpy as np
import pytest

from states import gr_state, sim_params

pytestmark = pytest.mark.gpu

AX = (-3, -2, -1)


def rel(a, b):
    d = np.max(np.abs(np.asarray(a) - np.asarray(b)))
    s = np.max(np.abs(b))
    return d / s if s > 0 else d


def new_box(gev, st, N, form):
    b = gev.Box(N, len(st["ids"]), form=form)
    b.upload_particles(st["pos"], st["q"], st["ids"], st["pmass"])
    for f, key in ((gev.PHI, "phi"), (gev.CHI, "chi"), (gev.BI, "Bi"), (gev.BIFT, "BiFT"),
                   (gev.SCALARFT, "scalarFT")):
        b.upload_field(f, st[key])
    return b


def sim_and_state(gev, sp, N):
    sim = gev.SimParams(sp["gr_flag"], sp["vector_flag"], sp["Cf"], sp["steplimit"], sp["z_relax"],
                        sp["fourpiG"], gev.make_cosmology())
    a0 = 1.0 / (1.0 + sp["z_in"])
    return sim, gev.CycleState(a0, 0.0, gev.next_dtau(sim, N, a0), 0.0, 0.0, 0.0, 0.0, 0, 0)


@pytest.mark.parametrize("N", [16, 24, 64, 96])
def test_fp32_fft_against_numpy(gev, N):
    rng = np.random.default_rng(N)
    x = rng.standard_normal((N, N, N))
    ref = np.fft.rfftn(x)
    with gev.Box(N, 8, form=gev.FORM_FP32) as b:
        b.upload_field(gev.SOURCE, x)
        b.fft(gev.PLAN_SOURCE, gev.FFT_FORWARD)
        k = b.download_field(gev.SCALARFT)
        assert rel(k, ref) <= 2e-6, rel(k, ref)
        b.fft(gev.PLAN_PHI, gev.FFT_BACKWARD)
        back = b.download_field(gev.PHI) / N ** 3
    x32 = x.astype(np.float32).astype(np.float64)
    assert rel(back, x32) <= 2e-6, rel(back, x32)


def test_fp32_kspace_solves_and_stencils_track_fp64(gev):
    """The standalone entry points in the FP32 form: each output within float
    rounding of the FP64 form's on the same (float-representable) inputs."""
    N = 32
    st = gr_state(N, 9)
    st = {k: (v.astype(np.float32).astype(np.float64) if k in ("phi", "chi", "Bi") else v) for k, v in st.items()}
    rng = np.random.default_rng(2)
    S = rng.standard_normal((6, N, N, N)).astype(np.float32).astype(np.float64)
    out = {}
    for form in (0, 1):
        b = new_box(gev, st, N, form)
        b.upload_field(gev.SIJ, S)
        b.upload_field(gev.SOURCE, S[0])
        b.prepareFTsource_scalar(0.31, 1e-3, 2e-4, 3e-5)
        b.prepareFTsource_tensor(2e-4)
        r = {"source": b.download_field(gev.SOURCE), "Sij": b.download_field(gev.SIJ)}
        b.fft(gev.PLAN_SIJ, gev.FFT_FORWARD)
        b.projectFTscalar(0)
        b.evolveFTvector(1e-3)
        r["scalarFT"] = b.download_field(gev.SCALARFT)
        r["BiFT"] = b.download_field(gev.BIFT)
        b.fft(gev.PLAN_SOURCE, gev.FFT_FORWARD)
        b.solveModifiedPoissonFT(1.0 / (1.0 / N) ** 2, 0.5)
        r["poisson"] = b.download_field(gev.SCALARFT)
        r["sum"] = b.field_sum(gev.PHI)
        out[form] = r
        b.close()
    for key in ("source", "Sij"):
        assert rel(out[1][key], out[0][key]) <= 1e-6, (key, rel(out[1][key], out[0][key]))
    for key in ("scalarFT", "BiFT", "poisson"):
        assert rel(out[1][key], out[0][key]) <= 1e-5, (key, rel(out[1][key], out[0][key]))
    # phi holds float-representable values: the FP64 sums agree to summation order
    assert abs(out[1]["sum"] - out[0]["sum"]) <= 1e-12 * np.abs(st["phi"]).sum()


def run_forms(gev, N, boxsize, ncycles, gr=True, elliptic=False, seed=31):
    st = gr_state(N, seed)
    sp = sim_params(N, boxsize, gr=gr, vector_elliptic=elliptic)
    sim, s64 = sim_and_state(gev, sp, N)
    _, s32 = sim_and_state(gev, sp, N)
    b64, b32 = new_box(gev, st, N, gev.FORM_FP64), new_box(gev, st, N, gev.FORM_FP32)
    q0 = st["q"]
    errs = []
    for c in range(ncycles):
        b64.run_cycles(sim, s64, 1)
        b32.run_cycles(sim, s32, 1)
        assert s64.a == s32.a and s64.dtau == s32.dtau  # FP64 background in both forms
        e = {}
        for name, f in (("phi", gev.PHI), ("chi", gev.CHI), ("Bi", gev.BI)):
            e[name] = rel(b32.download_field(f), b64.download_field(f))
        phimax = np.max(np.abs(b64.download_field(gev.PHI)))
        p64, q64, i64 = b64.download_particles()
        p32, q32, i32 = b32.download_particles()
        o64, o32 = np.argsort(i64), np.argsort(i32)
        assert np.array_equal(i64[o64], i32[o32])
        e["pos_abs"] = np.max(np.abs(p32[o32] - p64[o64]))
        e["dq"] = rel(q32[o32] - q0, q64[o64] - q0)
        e["T00hom"] = abs(s32.T00hom - s64.T00hom) / abs(s64.T00hom) if gr else 0.0
        # phi(k=0), the mean of phi, against the scale of phi
        e["phi_k0"] = abs(s32.phi_k0 - s64.phi_k0) / phimax
        errs.append(e)
    b64.close()
    b32.close()
    return errs


@pytest.mark.parametrize("N,gr,elliptic", [(32, True, False), (48, True, False), (32, True, True),
                                           (32, False, False)])
def test_fp32_cycles_track_fp64(gev, N, gr, elliptic):
    errs = run_forms(gev, N, 5.0 * N, 4, gr=gr, elliptic=elliptic)
    for c, e in enumerate(errs):
        print("cycle", c, {k: "%.2e" % v for k, v in e.items()})
        assert e["phi"] <= 1e-4 and e["chi"] <= 1e-4, (c, e)
        assert e["Bi"] <= 1e-3 or not gr, (c, e)
        assert e["pos_abs"] <= 1e-9, (c, e)
        assert e["dq"] <= 1e-5, (c, e)
        assert e["T00hom"] <= 1e-6, (c, e)
        assert e["phi_k0"] <= 1e-4, (c, e)


def test_fp32_run_to_run_bit_identical(gev):
    N = 32
    st = gr_state(N, 4)
    sp = sim_params(N, 160.0)
    outs = []
    for _ in range(2):
        sim, s = sim_and_state(gev, sp, N)
        b = new_box(gev, st, N, gev.FORM_FP32)
        b.run_cycles(sim, s, 3)
        outs.append([b.download_field(f) for f in (gev.PHI, gev.CHI, gev.BI, gev.BIFT, gev.SOURCE, gev.SIJ)]
                    + list(b.download_particles()) + [np.array([s.maxvel, s.T00hom, s.phi_k0])])
        b.close()
    for x, y in zip(*outs):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("gr,elliptic", [(True, False), (True, True), (False, False)])
def test_fp32_fused_cycle_equals_call_sequence(gev, gr, elliptic):
    from test_gpu_parity import cycle_by_calls
    N = 32
    st = gr_state(N, 5)
    sp = sim_params(N, 5.0 * N, gr=gr, vector_elliptic=elliptic)
    sim, sg = sim_and_state(gev, sp, N)
    fused, calls = new_box(gev, st, N, gev.FORM_FP32), new_box(gev, st, N, gev.FORM_FP32)
    sc = {"a": sg.a, "tau": 0.0, "dtau": sg.dtau, "dtau_old": 0.0}
    for cyc in range(3):
        fused.run_cycles(sim, sg, 1)
        cycle_by_calls(gev, calls, sp, sim, sc)
        for f in (gev.PHI, gev.CHI, gev.BI, gev.BIFT):
            x, y = fused.download_field(f), calls.download_field(f)
            assert np.array_equal(x, y), (cyc, f, rel(x, y))
        pf, qf, idf = fused.download_particles()
        pc, qc, idc = calls.download_particles()
        assert np.array_equal(pf, pc) and np.array_equal(qf, qc) and np.array_equal(idf, idc), cyc
    fused.close()
    calls.close()


def delta_spectrum(gev, b, a, numbins):
    """output.hpp:1990-2028: T00 with phi, forward FFT, extractPowerSpectrum."""
    b.projection_init(gev.SOURCE)
    b.projection_T00_project(a, True, 1.0)
    b.fft(gev.PLAN_SOURCE, gev.FFT_FORWARD)
    return b.extractPowerSpectrum(gev.SCALARFT, numbins)


def test_copy_state_and_lcdm_power_spectrum(gev):
    """Generated LCDM initial conditions (FP64 generator) carried into an FP32
    box with gev_copy_state, both forms evolved 6 cycles; the matter power
    spectrum of each, measured in its own form, agrees to 1e-4 below k_Nyq/2."""
    from test_gpu_ic import ic_inputs
    N, boxsize = 64, 320.0
    cosmo_o, fpg, tabs, tpl, sim_o, ic_o = ic_inputs(N, boxsize, 1)
    sim = gev.SimParams(1, 0, 48.0, 0.04, 30.0, fpg, gev.make_cosmology())
    ic = gev.ICParams(ic_o.seed, 1, 1, N // 4, 100.0, 2.1e-9, 0.965, 0.05, boxsize)
    b64 = gev.Box(N, N ** 3)
    b32 = gev.Box(N, N ** 3, form=gev.FORM_FP32)
    with pytest.raises(gev.GevError):
        b32.generateIC_basic(sim, ic, tpl, *tabs)  # the generator needs an FP64 box
    mv, _ = b64.generateIC_basic(sim, ic, tpl, *tabs)
    b32.copy_state_from(b64)
    # the carried state: particles exact, fields rounded to float
    p64, q64, i64 = b64.download_particles()
    p32, q32, i32 = b32.download_particles()
    assert np.array_equal(p64, p32) and np.array_equal(q64, q32) and np.array_equal(i64, i32)
    for f in (gev.PHI, gev.BI, gev.BIFT):
        x = b64.download_field(f)
        x32 = x.astype(np.complex64 if np.iscomplexobj(x) else np.float32).astype(x.dtype)
        assert np.array_equal(b32.download_field(f), x32), f
    assert np.array_equal(b32.download_filing()[0], b64.download_filing()[0])
    # an FP64 copy of the FP64 state is exact
    with gev.Box(N, N ** 3) as c64:
        c64.copy_state_from(b64)
        for f in (gev.PHI, gev.CHI, gev.BI, gev.BIFT):
            assert np.array_equal(c64.download_field(f), b64.download_field(f))

    a0 = 1.0 / 101.0
    s64 = gev.CycleState(a0, 0.0, gev.next_dtau(sim, N, a0), 0.0, mv, 0.0, 0.0, 0, 0)
    s32 = gev.CycleState(a0, 0.0, gev.next_dtau(sim, N, a0), 0.0, mv, 0.0, 0.0, 0, 0)
    b64.run_cycles(sim, s64, 6)
    b32.run_cycles(sim, s32, 6)
    assert s64.a == s32.a
    nb = N // 2
    g64 = delta_spectrum(gev, b64, s64.a, nb)
    g32 = delta_spectrum(gev, b32, s32.a, nb)
    assert np.array_equal(g64[4], g32[4])
    kny = np.pi * N  # k in units of 2 pi / box: the Nyquist wave number of the grid
    low = (g64[4] > 0) & (g64[0] < 0.5 * kny)
    err = np.max(np.abs(g32[1][low] - g64[1][low]) / np.abs(g64[1][low]))
    print("P(k) max relative error below k_Nyq/2: %.2e over %d bins" % (err, int(low.sum())))
    assert low.sum() >= 8
    assert err <= 1e-4, err
    b64.close()
    b32.close()


def test_fp32_dense_tiles(gev):
    """Tiles whose source region holds more particles than the LDS records
    (a cell of 2500 particles, cells of 100 and 20) take the fixed-point
    particle-parallel deposit in the FP32 form: within float rounding of the
    FP64 deposit, bit-reproducible, and the same for the fused T00+Tij call
    as for the two separate calls."""
    N = 16
    rng = np.random.default_rng(5)
    groups = [(3, 0.41), (20, 0.27), (100, 0.66), (2500, 0.09), (1, 0.8)]
    pos = [np.tile([[c0, c0 + 0.01, c0 + 0.02]], (n, 1)) + rng.uniform(0, 0.05 / N, size=(n, 3))
           for n, c0 in groups]
    pos = np.concatenate(pos + [rng.uniform(0, 1, size=(300, 3))])
    n = len(pos)
    ids = rng.permutation(np.arange(10 * n, dtype=np.int64))[:n]
    q = rng.standard_normal((n, 3)) * 1e-3
    phi = rng.standard_normal((N, N, N)) * 1e-5
    a = 0.3
    res = {}
    for form in (gev.FORM_FP64, gev.FORM_FP32, gev.FORM_FP32):
        with gev.Box(N, n, form=form) as b:
            b.upload_particles(pos, q, ids, 1e-3)
            b.upload_field(gev.PHI, phi)
            b.projection_init(gev.SOURCE)
            b.projection_init(gev.SIJ)
            b.projection_T00_Tij_project(a, 1.0)
            fused = (b.download_field(gev.SOURCE), b.download_field(gev.SIJ))
            b.projection_init(gev.SOURCE)
            b.projection_T00_project(a, True, 1.0)
            b.projection_init(gev.SIJ)
            b.projection_Tij_project(a, True, 1.0)
            sep = (b.download_field(gev.SOURCE), b.download_field(gev.SIJ))
        res.setdefault(form, []).append((fused, sep))
    (f64, _), = res[gev.FORM_FP64]
    (f32a, s32a), (f32b, _) = res[gev.FORM_FP32]
    for x, y in zip(f32a, f32b):
        assert np.array_equal(x, y)  # run to run
    for x, y in zip(f32a, s32a):
        assert np.array_equal(x, y)  # fused = separate
    assert rel(f32a[0], f64[0]) <= 1e-6, rel(f32a[0], f64[0])
    for c in range(6):
        assert rel(f32a[1][c], f64[1][c]) <= 1e-5, (c, rel(f32a[1][c], f64[1][c]))


@pytest.mark.parametrize("N", [24, 32])
def test_fp32_tiles_above_the_records(gev, N):
    """FP32 tiles whose source region holds more particles than the LDS
    records take one of two paths: the chunked gather when every source cell
    is small (2 particles a cell: 850 in a region), the fixed-point path when
    a cell is dense (a cell of 40 on top of that; its records spread over the
    lanes) or the region is crowded (4 a cell in one corner of the box).
    Each within float rounding of the FP64 deposit, bit-reproducible.  At
    N = 24 the x tiles of 16 and 8 cells take the general (not full-tile)
    paths."""
    rng = np.random.default_rng(11)
    g = (np.stack(np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij"), -1).reshape(-1, 3) + 0.5) / N
    pos = [np.repeat(g, 2, axis=0) + rng.uniform(-0.45, 0.45, size=(2 * N ** 3, 3)) / N]
    corner = g[np.all(g < 0.25, axis=1)]
    pos.append(np.repeat(corner, 2, axis=0) + rng.uniform(-0.45, 0.45, size=(2 * len(corner), 3)) / N)
    pos.append(np.array([[0.6, 0.6, 0.6]]) + rng.uniform(0.001, 0.03, size=(40, 3)) / N)
    pos = np.mod(np.concatenate(pos), 1.0)
    n = len(pos)
    ids = rng.permutation(np.arange(n, dtype=np.int64))
    q = rng.standard_normal((n, 3)) * 1e-3
    phi = rng.standard_normal((N, N, N)) * 1e-5
    out = []
    for form in (gev.FORM_FP64, gev.FORM_FP32, gev.FORM_FP32):
        with gev.Box(N, n, form=form) as b:
            b.upload_particles(pos, q, ids, 1e-3)
            b.upload_field(gev.PHI, phi)
            b.projection_init(gev.SOURCE)
            b.projection_init(gev.SIJ)
            b.projection_T00_Tij_project(0.4, 1.0)
            out.append((b.download_field(gev.SOURCE), b.download_field(gev.SIJ)))
    f64, f32a, f32b = out
    for x, y in zip(f32a, f32b):
        assert np.array_equal(x, y)  # run to run
    assert rel(f32a[0], f64[0]) <= 1e-6, rel(f32a[0], f64[0])
    for c in range(6):
        assert rel(f32a[1][c], f64[1][c]) <= 1e-5, (c, rel(f32a[1][c], f64[1][c]))


def test_fp32_form_arguments(gev):
    with pytest.raises(gev.GevError):
        gev.Box(32, 8, form=7)
    with gev.Box(32, 8) as a, gev.Box(16, 8, form=gev.FORM_FP32) as b:
        with pytest.raises(gev.GevError):
            b.copy_state_from(a)  # grid sizes differ
