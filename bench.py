"""Throughput of the GR particle-mesh cycle on one MI355X.

Workload (the plain FP64 GPU drop-in configuration): GR gravity, 512^3 grid,
512^3 CDM particles, box 2048 Mpc/h, h = 0.67, Omega_cdm = 0.26, Omega_b =
0.05 folded into CDM, Courant factor 48, time step limit 0.04, vector method
parabolic, FP64, z_in = 100.  A "step" is one cycle of gevolution's main
loop (main.cpp:372-879): T00+Tij deposit, both prepareFTsource stencils, the
12 FFTs with the k-space solves, and the fused kick + drift + re-filing.

Initial state: the basic IC generator on the device (generateIC_basic of
ic_basic.hpp) from a BBKS transfer table in CLASS format (bbks.py; h = 0.67,
Omega_m = 0.31, A_s = 2.1e-9, n_s = 0.965, k_pivot = 0.05/Mpc), the
reference's sc1 crystal template tiled N/4 times, Gaussian modes from seed 7,
k-domain sphere, template-corrected displacement.  The IC generation itself
is timed once and reported beside the cycle throughput ("ic_generator").
`--state lattice` instead starts from the bare cell-centred lattice with a
+-0.05-cell jitter and small random momenta (the CPU baseline's sample).

Output: one JSON line.  `value` is particle-steps/s over the timed cycles
with everything resident in HBM.  The timed cycles start from the initial
conditions at z = 100 (the warm-up cycles run first and the initial
conditions are then generated again); the default 117 cycles are the whole
run to z = 0.  `roofline`
is the kernel with the largest share of the timed region, its algorithmic
bytes per launch divided by its mean duration (HIP events on the context
stream); `cpu_baseline` is the CPU oracle (the restated reference
algorithm, OpenMP) on a bounded sample of the same workload;
`precision_form` times the FP32 form over the same cycles from the same
state; `full_run` runs the whole configuration, z = 100 to z = 0 (117
cycles), in both forms: the time to z = 0 and the FP32 form's errors at
z = 10, 1 and 0; `serving` serves a Poisson trace of whole simulations
(64^3 to 256^3, GR and Newtonian, z = 100 to 0) with overlapping requests
and reused boxes, against back-to-back plain runs of the same trace: p50 /
p99 time to completion and aggregate particle-steps/s of both.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gevolution-1.2_amd"))

METRIC = "GR PM particle-steps/s at 512^3 (%HBM peak); p99 ensemble latency vs plain"
HBM_PEAK_GBS = 8000.0
C_LIGHT = 2997.92458


def algorithmic_bytes(name, N, Np, f=8):
    """Bytes each kernel must move at least once per launch: every input read
    once, every output written once.  f: bytes of a real field element (8 in
    the FP64 form, 4 in the FP32 form; particles are FP64 in both)."""
    n3 = N ** 3
    nk = N * N * (N // 2 + 1)
    c = 2 * f  # a complex k-field element
    table = {
        "kick_drift": 96 * Np + 5 * f * n3,                 # pos,q r+w; phi, chi, B(3) read
        "update_vel": 72 * Np + 5 * f * n3,
        "move_particles": 72 * Np + 5 * f * n3,
        "projection_T00_Tij_project": 48 * Np + 8 * f * n3,  # pos,q read; phi read; T00 + Tij(6) written
        "projection_T00_project": 48 * Np + 2 * f * n3,
        "projection_Tij_project": 48 * Np + 7 * f * n3,
        "prepare_ft_source_tensor": 13 * f * n3,            # phi + Tij(6) read, Sij(6) written
        "prepare_ft_source_scalar": 4 * f * n3,             # phi, chi, source read, source written
        "fft_r2c_x": f * n3 + c * nk,
        "fft_c2r_x": c * nk + f * n3,
        "fft_c2c_y": 2 * c * nk,
        "fft_c2c_z": 2 * c * nk,
        "fft_c2c_y_inv": 2 * c * nk,
        "fft_c2c_z_inv": 2 * c * nk,
        "solve_modified_poisson_ft": 2 * c * nk,
        "project_scalar_evolve_vector": 13 * c * nk,        # SijFT(6)+BiFT(3) read, chiFT+BiFT(3) written
        # fused passes (fftsolve.hip)
        "prepare_tensor_fft_r2c_x": 7 * f * n3 + 6 * c * nk,   # phi + Tij(6) read, 6 k rows written
        "prepare_scalar_fft_r2c_x": 3 * f * n3 + c * nk,       # phi, chi, source read, 1 k row written
        "fft_z_poisson_z": 2 * c * nk,                         # y-passed source in, z-inverted phi out
        "fft_z_tensor_z": 16 * c * nk,                         # Sij(6)+BiFT(3) in, BiFT(3)+chi+B(3) out
        "fft_z_tensor_scalar_z": 7 * c * nk,
        "fft_z_vector_z": 9 * c * nk,
        "field_sum": f * n3,
        "filing_scan": 8 * n3,
        "filing_scatter": 12 * Np,
        "filing_sort": 4 * n3,
    }
    # a pass launched over n components at once is timed as "name[n]"
    if name.endswith("]") and "[" in name:
        base, n = name[:-1].split("[")
        b = table.get(base)
        return None if b is None else b * int(n)
    return table.get(name)


# Algorithmic bytes of one GR-parabolic cycle at the reference's stage
# boundaries (FP64 fields and particles): each stage reads its inputs once and
# writes its outputs once, an FFT counts one read and one write of its whole
# array (c = bytes per cell of one complex half-spectrum component):
#   deposit T00 + Tij        48 Np + 64 N^3
#   prepare scalar + T00hom  32 N^3
#   FFT source + Poisson     2c N^3,  FFT -> phi  2c N^3
#   prepare tensor           104 N^3
#   FFT Sij x6               12c N^3
#   projectFTscalar + evolveFTvector  13c N^3
#   FFT -> chi 2c N^3,  FFT -> B x3  6c N^3
#   kick + drift             96 Np + 40 N^3
#   cell key + histogram     8 N^3
# (f = 8; in the FP32 form f = 4 for every field term, the cell key stays 8 N^3)
def stage_bytes(N, Np, f=8):
    c = f * (1.0 + 2.0 / N)
    return int(144 * Np + (30 * f + 8) * N ** 3 + 37 * c * N ** 3)


# Minimum bytes of one GR-parabolic cycle (FP64 fields and particles): every
# carried field and particle array read once and written once per stage that
# has to materialise it; the FFT chains count as one read of their inputs and
# one write of their outputs (no pass intermediates, no k-space copies except
# the carried BiFT).
def cycle_bytes(N, Np):
    n3 = N ** 3
    nk = N * N * (N // 2 + 1)
    return (48 * Np + 8 * n3 + 56 * n3  # deposit: pos, q, phi in; T00 + Tij out
            + 8 * n3                    # T00hom: source re-read
            + 24 * n3 + 8 * n3          # scalar chain: phi, chi, source in; phi out
            + 56 * n3 + 48 * nk         # tensor chain: phi, Tij in; BiFT in
            + 48 * nk + 32 * n3         # ... BiFT, chi, B out
            + 96 * Np + 40 * n3         # kick + drift: pos, q in and out; phi, chi, B in
            + 8 * Np + 4 * n3           # cell + slot out; counts
            + 8 * n3 + 12 * Np)         # CSR scan; perm scatter


def measured_traffic(path, kernel):
    """HBM bytes per launch of `kernel` from the rocprofv3 --pmc summary
    (lines 'kernel read_bytes written_bytes traffic_bytes'), or None."""
    try:
        for line in open(path):
            parts = line.split()
            if len(parts) == 4 and parts[0] == kernel and parts[3].isdigit():
                return int(parts[3])
    except OSError:
        pass
    return None


def synthetic_state(N, seed):
    rng = np.random.default_rng(seed)
    Np = N ** 3
    g = (np.arange(N, dtype=np.float64) + 0.5) / N
    pos = np.empty((Np, 3))
    pos[:, 0] = np.tile(g, N * N)
    pos[:, 1] = np.tile(np.repeat(g, N), N)
    pos[:, 2] = np.repeat(g, N * N)
    pos += rng.uniform(-0.05 / N, 0.05 / N, size=pos.shape)
    q = rng.standard_normal(size=(Np, 3)) * 2e-6
    ids = np.arange(Np, dtype=np.int64)
    return pos, q, ids


def ic_state(gev, box, N, boxsize, seed):
    """Runs the device IC generator; returns (sim, state, ic_ms, maxvel, max_disp)."""
    import bbks
    h = 0.67
    k, d, t = bbks.bbks_table(np.logspace(-5.0, np.log10(300.0), 400), 100.0, 0.31, h)
    kb, tb = k * boxsize, t * boxsize / h  # loadTransferFunctions units
    tpl = gev.load_homogeneous_template(os.path.join(ROOT, "tests", "golden", "sc1_crystal.dat"))
    sim, st = sim_setup(gev, N, boxsize)
    ic = gev.ICParams(seed, 1, 1, N // 4, 100.0, 2.1e-9, 0.965, 0.05, boxsize)
    box.synchronize()
    t0 = time.perf_counter()
    mv, md = box.generateIC_basic(sim, ic, tpl, kb, d, tb, d, tb)
    box.synchronize()
    ic_ms = 1e3 * (time.perf_counter() - t0)
    st.maxvel = mv
    return sim, st, ic_ms, mv, md


def sim_setup(gev, N, boxsize):
    cosmo = gev.make_cosmology(0.26, 0.05, 0.67)
    fourpiG = 1.5 * boxsize * boxsize / C_LIGHT / C_LIGHT
    sim = gev.SimParams(1, 0, 48.0, 0.04, 100.0, fourpiG, cosmo)
    a0 = 1.0 / 101.0
    st = gev.CycleState(a0, 0.0, gev.next_dtau(sim, N, a0), 0.0, 0.0, 0.0, 0.0, 0, 0)
    return sim, st


def hbm_probe(gev):
    """The device's achievable HBM bandwidth: the library's hand-written
    streaming kernels over 2 GiB buffers (copy, copy with non-temporal hints,
    read, write), GB/s and the fraction of the 8 TB/s peak."""
    out = {}
    for name, kind in (("copy", gev.PROBE_COPY), ("copy_nt", gev.PROBE_COPY_NT), ("read", gev.PROBE_READ),
                       ("write", gev.PROBE_WRITE)):
        gbs, ms = gev.bandwidth_probe(2 << 30, kind, 20)
        out[name] = {"GBps": round(gbs, 1), "ms": round(ms, 4), "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def host_info(threads):
    """The host the CPU baseline ran on: cores seen, CPU model, threads used
    and a STREAM-style triad bandwidth with those threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    orc.build()
    orc.set_threads(threads)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model, "threads": threads,
            "stream_triad_gbs": round(orc.stream_triad_gbs(), 1)}


def cpu_baseline(state, N, boxsize, threads):
    """Times the CPU oracle (the FP64 C restatement of the reference
    algorithm, OpenMP) on a bounded sample of the workload: the device's
    generated initial state at N^3, cycle 0 untimed (it skips the phi
    update), then cycle 1 timed, with main.cpp's BENCHMARK categories."""
    import oracle as orc
    orc.set_threads(threads)
    ob = orc.OracleBox(N, N ** 3)
    order = np.argsort(state["ids"])
    ob.pos[:] = state["pos"][order]
    ob.q[:] = state["q"][order]
    ob.id[:] = state["ids"][order]
    del order
    ob.box.pmass = state["pmass"]
    for name in ("phi", "chi", "Bi", "BiFT", "scalarFT"):
        getattr(ob, name)[:] = state[name]
    ob.file()
    sim = orc.Sim(N, 1, 0, 48.0, 0.04, 100.0, 1.5 * boxsize ** 2 / C_LIGHT ** 2, orc.make_cosmo())
    a0 = 1.0 / 101.0
    st = orc.State(a0, 0.0, orc.lib().orc_next_dtau(orc.ctypes.byref(sim), a0), 0.0, 0.0, 0.0, 0.0, 0)
    ob.cycle(sim, st)
    orc.timers_reset()
    t0 = time.perf_counter()
    ob.cycle(sim, st)
    dt = time.perf_counter() - t0
    cats = orc.timers()
    ob.close()
    return N ** 3 / dt, dt, {k: round(v[0], 3) for k, v in cats.items() if k != "spectra_output"}


def cpu_c1_full(threads):
    """The plain CPU configuration in full on the oracle: GR, 64^3 grid and
    particles, box 320 Mpc/h, BBKS LCDM ICs from seed 42, z = 100 to 0 with
    the delta spectrum at z = 0; particle-steps/s over the main loop and the
    BENCHMARK categories."""
    import bbks
    import oracle as orc
    orc.set_threads(threads)
    N, L, h = 64, 320.0, 0.67
    k, d, t = bbks.bbks_table(np.logspace(-5.0, np.log10(300.0), 400), 100.0, 0.31, h)
    sim = orc.Sim(N, 1, 0, 48.0, 0.04, 100.0, 1.5 * L * L / C_LIGHT ** 2, orc.make_cosmo())
    ob = orc.OracleBox(N, N ** 3)
    tpl = orc.load_template(os.path.join(ROOT, "tests", "golden", "sc1_crystal.dat"))
    ob.generate_ic(sim, orc.IC(42, 1, 1, N // 4, 1, 100.0, 2.1e-9, 0.965, 0.05, L), tpl,
                   k * L, d, t * L / h, d, t * L / h)
    a0 = 1.0 / 101.0
    st = orc.State(a0, orc.particle_horizon(a0, sim.fourpiG, sim.cosmo),
                   orc.lib().orc_next_dtau(orc.ctypes.byref(sim), a0), 0.0, 0.0, 0.0, 0.0, 0)
    orc.timers_reset()
    t0 = time.perf_counter()
    ob.run_to_outputs(sim, st, [0.0], N // 2, L)
    dt = time.perf_counter() - t0
    cats = orc.timers()
    ob.close()
    return {"config": "GR 64^3, box 320 Mpc/h, BBKS seed 42, z = 100 -> 0, P(k) at z = 0",
            "cycles": st.cycle, "main_loop_s": round(dt, 3), "value": round(N ** 3 * st.cycle / dt, 1),
            "unit": "particle-steps/s", "categories_s": {k: round(v[0], 3) for k, v in cats.items()}}


def download_state(gev, box):
    """The box's carried state on the host (the CPU baseline's sample)."""
    pos, q, ids = box.download_particles()
    st = {"pos": pos, "q": q, "ids": ids, "pmass": 0.31 / len(ids)}
    for f, name in ((gev.PHI, "phi"), (gev.CHI, "chi"), (gev.BI, "Bi"), (gev.BIFT, "BiFT"), (gev.SCALARFT, "scalarFT")):
        st[name] = box.download_field(f)
    return st


def kernel_breakdown(kernels, steps, N, Np, f):
    """Per-kernel calls per step, mean duration, share and achieved GB/s."""
    out = {}
    total_ms = sum(ms for _, ms in kernels.values())
    for name, (calls, ms) in sorted(kernels.items(), key=lambda kv: -kv[1][1]):
        avg = ms / calls
        b = algorithmic_bytes(name, N, Np, f)
        out[name] = {"calls_per_step": calls / steps, "ms_per_launch": round(avg, 4),
                     "share": round(ms / total_ms, 4),
                     "GBps": round(b / (avg * 1e-3) / 1e9, 1) if b else None}
    return out


def timed_cycles(box, sim, st, warmup, steps, timers):
    """warmup untimed cycles, then steps timed ones; (seconds, per-kernel table)."""
    box.run_cycles(sim, st, warmup)
    box.synchronize()
    if timers:
        box.timing_reset()
        box.timing(True)
    t0 = time.perf_counter()
    box.run_cycles(sim, st, steps)
    box.synchronize()
    elapsed = time.perf_counter() - t0
    kernels = box.timing_report() if timers else {}
    box.timing(False)
    return elapsed, kernels


def delta_spectrum(gev, box, a, numbins):
    """The matter power spectrum of output.hpp:1990-2028: T00 with phi, FFT,
    extractPowerSpectrum (deconvolved, linear k in units of 2 pi / box)."""
    box.projection_init(gev.SOURCE)
    box.projection_T00_project(a, True, 1.0)
    box.fft(gev.PLAN_SOURCE, gev.FFT_FORWARD)
    return box.extractPowerSpectrum(gev.SCALARFT, numbins)


def precision_form_report(gev, box, st, box32, st32, sim, args, ms64):
    """Times the FP32 form over the same cycles from the same initial state
    and states its error against the FP64 form: phi (max norm) and the
    matter power spectrum below k_Nyq/2."""
    N = box.N
    el, k32 = timed_cycles(box32, sim, st32, 0, args.steps, not args.no_timers)
    ms32 = 1e3 * el / args.steps
    assert st32.a == st.a
    phi64 = box.download_field(gev.PHI)
    phi32 = box32.download_field(gev.PHI)
    phi_err = float(np.max(np.abs(phi32 - phi64)) / np.max(np.abs(phi64)))
    del phi64, phi32
    nb = N // 2
    p64 = delta_spectrum(gev, box, st.a, nb)
    p32 = delta_spectrum(gev, box32, st32.a, nb)
    low = (p64[4] > 0) & (p64[0] < 0.5 * np.pi * N)
    pk_err = float(np.max(np.abs(p32[1][low] - p64[1][low]) / np.abs(p64[1][low])))
    print("FP32 form: %.3f ms/cycle (%.2fx), phi error %.2e, P(k) error %.2e"
          % (ms32, ms64 / ms32, phi_err, pk_err), file=sys.stderr, flush=True)
    return {"form": "fp32 fields, FFT passes, deposit sums and field stencils; fp64 particles, background, "
                    "momentum factors and the k-space solves' coefficients and denominators (k->0, zero mode)",
            "ms_per_step": round(ms32, 3), "value": round(box.N ** 3 * args.steps / el, 1),
            "speedup_vs_fp64": round(ms64 / ms32, 3),
            "phi_max_rel_err": phi_err, "pk_max_rel_err_below_half_nyquist": pk_err,
            "pk_bins_compared": int(low.sum()), "cycles_from_same_state": args.steps,
            "z_at_comparison": round(1.0 / st.a - 1.0, 4),
            "step_bytes": stage_bytes(N, N ** 3, 4),
            "kernels": kernel_breakdown(k32, args.steps, N, N ** 3, 4) if k32 else {}}


def full_run_report(gev, box, box32, N, boxsize, seed):
    """The whole run of the workload: generated ICs at z = 100 evolved to
    z = 0 in both forms through the main loop with the delta power spectrum
    written at z = 10, 1 and 0 as the reference writes it
    (gev_run_to_outputs: EXACT_OUTPUT_REDSHIFTS interpolation, the loop
    stopping after the field stage of the z = 0 cycle).  Each leg is
    synchronised and timed; the comparisons are outside the timings.  Late
    cycles cost more than early ones (the clustered particles make the
    deposit and the filing slower), so this is the time a user of the
    configuration waits.  The FP32 form's phi and P(k) errors against the
    FP64 form are stated at z = 10, 1 and 0."""
    sim, st, _, _, _ = ic_state(gev, box, N, boxsize, seed)
    st32 = None
    if box32 is not None:
        box32.copy_state_from(box)
        st32 = gev.CycleState(st.a, st.tau, st.dtau, st.dtau_old, st.maxvel, st.T00hom, st.phi_k0, st.cycle, 0)
    t64 = t32 = 0.0
    points = []
    for zt in (10.0, 1.0, 0.0):
        box.synchronize()
        t0 = time.perf_counter()
        o64 = box.run_to_outputs(sim, st, [zt], N // 2, boxsize)
        t64 += time.perf_counter() - t0
        pt = {"z": float(o64["z"][0]), "cycle": st.cycle, "fp64_s": round(t64, 3)}
        if box32 is not None:
            t0 = time.perf_counter()
            o32 = box32.run_to_outputs(sim, st32, [zt], N // 2, boxsize)
            t32 += time.perf_counter() - t0
            assert st32.a == st.a and st32.cycle == st.cycle
            pt["fp32_s"] = round(t32, 3)
            phi64 = box.download_field(gev.PHI)
            phi32 = box32.download_field(gev.PHI)
            pt["phi_max_rel_err"] = float(np.max(np.abs(phi32 - phi64)) / np.max(np.abs(phi64)))
            del phi64, phi32
            low = (o64["occupation"][0] > 0) & (o64["k"][0] < 0.5 * np.pi * N / boxsize)
            pt["pk_max_rel_err_below_half_nyquist"] = float(np.max(np.abs(o32["pk"][0][low] - o64["pk"][0][low]) /
                                                                   np.abs(o64["pk"][0][low])))
        points.append(pt)
        print("full run: z = %.3f after %d cycles, %.2f s FP64%s" % (pt["z"], pt["cycle"], t64,
              (", %.2f s FP32" % t32) if box32 is not None else ""), file=sys.stderr, flush=True)
    cyc = st.cycle
    out = {"from_z": 100.0, "cycles": cyc, "time_to_z0_s": round(t64, 3),
           "ms_per_step": round(1e3 * t64 / cyc, 3), "value": round(N ** 3 * cyc / t64, 1),
           "unit": "particle-steps/s", "outputs": "delta P(k) at z = 10, 1, 0 (EXACT_OUTPUT_REDSHIFTS)",
           "checkpoints": points}
    if box32 is not None:
        out["fp32"] = {"time_to_z0_s": round(t32, 3), "ms_per_step": round(1e3 * t32 / cyc, 3),
                       "value": round(N ** 3 * cyc / t32, 1), "speedup_vs_fp64": round(t64 / t32, 3)}
    return out


def c3_report(gev):
    """The precision form at 1024^3 (box 4096 Mpc/h, BBKS LCDM ICs from seed
    11, GR): the FP64 form (compact: 277 GB, one particle set) and the FP32
    form (compact, 177 GB) from the same generated initial state, each run
    from z = 100 through the delta spectrum outputs at z = 10, 1 and 0
    (gev_run_to_outputs).  The two boxes do not fit the GPU together, so the
    state goes through host memory.  The FP32 form's phi (max norm) and P(k)
    (bins below k_Nyq/2) errors against the FP64 form at each output."""
    N, L, seed = 1024, 4096.0, 11
    Np = N ** 3
    zs = (10.0, 1.0, 0.0)
    box = gev.Box(N, Np, compact=True)
    sim, st, ic_ms, _, _ = ic_state(gev, box, N, L, seed)
    state = download_state(gev, box)
    print("c3: 1024^3 ICs in %.0f ms, state on the host" % ic_ms, file=sys.stderr, flush=True)
    ref = []
    t64 = 0.0
    for zt in zs:
        box.synchronize()
        t0 = time.perf_counter()
        o = box.run_to_outputs(sim, st, [zt], N // 2, L)
        t64 += time.perf_counter() - t0
        ref.append((st.cycle, o, box.download_field(gev.PHI)))
        print("c3: FP64 at z = %g after %d cycles, %.1f s" % (zt, st.cycle, t64), file=sys.stderr, flush=True)
    cyc = st.cycle
    box.close()
    b32 = gev.Box(N, Np, form=gev.FORM_FP32, compact=True)
    b32.upload_particles(state["pos"], state["q"], state["ids"], state["pmass"])
    for f, name in ((gev.PHI, "phi"), (gev.CHI, "chi"), (gev.BI, "Bi"), (gev.BIFT, "BiFT"), (gev.SCALARFT, "scalarFT")):
        b32.upload_field(f, state[name])
    del state
    st32 = sim_setup(gev, N, L)[1]
    t32, points = 0.0, []
    for zt, (c64, o64, phi64) in zip(zs, ref):
        b32.synchronize()
        t0 = time.perf_counter()
        o32 = b32.run_to_outputs(sim, st32, [zt], N // 2, L)
        t32 += time.perf_counter() - t0
        assert st32.cycle == c64
        phi32 = b32.download_field(gev.PHI)
        low = (o64["occupation"][0] > 0) & (o64["k"][0] < 0.5 * np.pi * N / L)
        points.append({"z": zt, "cycle": c64,
                       "phi_max_rel_err": float(np.max(np.abs(phi32 - phi64)) / np.max(np.abs(phi64))),
                       "pk_max_rel_err_below_half_nyquist": float(np.max(np.abs(o32["pk"][0][low] - o64["pk"][0][low])
                                                                         / np.abs(o64["pk"][0][low])))})
        del phi32
        print("c3: FP32 at z = %g, phi err %.2e, P(k) err %.2e" % (zt, points[-1]["phi_max_rel_err"],
              points[-1]["pk_max_rel_err_below_half_nyquist"]), file=sys.stderr, flush=True)
    b32.close()
    return {"config": "GR 1024^3 grid and particles, box 4096 Mpc/h, BBKS LCDM ICs seed 11, z = 100 -> 0",
            "boxes": "compact (one particle set): FP64 %.1f GB, FP32 %.1f GB"
                     % (gev.footprint_bytes(N, Np, gev.FORM_FP64, True) / 1e9,
                        gev.footprint_bytes(N, Np, gev.FORM_FP32, True) / 1e9),
            "cycles": cyc, "fp64_time_to_z0_s": round(t64, 2), "fp32_time_to_z0_s": round(t32, 2),
            "fp64_value": round(Np * cyc / t64, 1), "fp32_value": round(Np * cyc / t32, 1),
            "speedup_vs_fp64": round(t64 / t32, 3), "checkpoints": points}


def serving_report(args):
    """Whole simulations served on one GPU (serving.py) against back-to-back
    plain FP64 runs on the same Poisson arrival trace: request grids from
    {64, 96, 128, 192, 256}^3 (weights .3/.25/.2/.15/.1), half GR and half
    Newtonian, Omega_m and A_s drawn per request, each generated on the
    device and run from z = 100 to z = 0 with its delta spectrum at z = 0.
    The plain runs take one request at a time on a box of their own; the
    served runs are continuously batched (BatchServer: requests of one
    shape share a batch, join it at the next cycle and leave it when they
    finish).  Every request's spectrum and phi sum must be identical in
    both."""
    import gevpm as gev
    import serving as sv
    tpl = gev.load_homogeneous_template(os.path.join(ROOT, "tests", "golden", "sc1_crystal.dat"))
    be = sv.GevBackend(template=tpl)
    bbe = sv.BatchBackend(template=tpl)
    rates = [float(v) for v in args.serving_rates.split(",")]
    traces = {rate: sv.poisson_trace(args.serving_requests, rate, 2024) for rate in rates}
    # first use of each shape's kernels, allocations and graphs, outside both timings
    shapes = sorted({(r.ngrid, r.gr) for t in traces.values() for r in t})
    warm = [sv.Request(-1 - i, g, gr=gr) for i, (g, gr) in enumerate(shapes)]
    with sv.Server(be, sv.default_budget(), slots=1, reuse=False) as s:
        s.serve(warm)
    with sv.BatchServer(bbe, sv.default_budget()) as s:
        s.serve(warm)
    runs = []
    for rate, trace in traces.items():
        with sv.Server(be, sv.default_budget(), slots=1, reuse=False) as s:
            plain = s.serve(trace)
        with sv.BatchServer(bbe, sv.default_budget()) as s:
            served = s.serve(trace)
            pool = {"batches_created": s.created, "batches_evicted": s.evicted, "peak_running": s.peak_running,
                    "peak_device_bytes": s.peak_in_use, "rounds": s.rounds}
        same = all(a.status == b.status == "ok" and a.result["cycles"] == b.result["cycles"]
                   and np.array_equal(a.result["pk"], b.result["pk"]) and a.result["phi_sum"] == b.result["phi_sum"]
                   for a, b in zip(plain, served))
        ps, pp = sv.summarize(served), sv.summarize(plain)
        print("serving at %.0f/s: p99 %.3f s served, %.3f s plain; %.3e vs %.3e particle-steps/s"
              % (rate, ps["p99_s"], pp["p99_s"], ps["particle_steps_per_s"] or 0, pp["particle_steps_per_s"] or 0),
              file=sys.stderr, flush=True)
        runs.append({"rate_per_s": rate, "served": ps, "back_to_back_plain": pp,
                     "p99_speedup_vs_plain": round(pp["p99_s"] / ps["p99_s"], 3) if ps["p99_s"] else None,
                     "throughput_ratio_vs_plain": (round(ps["particle_steps_per_s"] / pp["particle_steps_per_s"], 3)
                                                   if ps["particle_steps_per_s"] and pp["particle_steps_per_s"]
                                                   else None),
                     "identical_results_alone_and_in_the_mix": bool(same), "pool": pool})
    return {"trace": {"requests": args.serving_requests, "seed": 2024,
                      "grids": [64, 96, 128, 192, 256], "weights": [0.3, 0.25, 0.2, 0.15, 0.1],
                      "gr_fraction": 0.5, "z": "100 -> 0", "form": "fp64"},
            "server": "continuous batching (one batch per grid and gravity, slots 8/4/2 by grid)",
            "plain": "back-to-back: one request at a time on a box of its own (gev_run_to_outputs)",
            "runs": runs}


def ensemble_report(args):
    """The batched ensemble: `ensemble_boxes` boxes with grids drawn
    uniformly from {32, 48, 64, 96, 128}^3 (particles = grid^3, box 5 Mpc/h
    per cell), Omega_m in [0.25, 0.35], A_s in [1.8, 2.4]e-9, seeds 0..n-1,
    GR, the FP32 precision form, z = 100 to 0 with the delta spectrum at
    z = 0.  The boxes of one grid size form one batch (one launch per kernel
    over all of them, both stages of the cycle replayed from HIP graphs), the
    five batches on their own streams.  Against the same boxes run one after
    another alone (gev_run_to_outputs, same form); a sample of boxes is
    checked bit for bit between the two, and the first box of each grid size
    is run in the FP64 form too: its z = 0 P(k) error."""
    import gevpm as gev
    import ensemble as en
    tpl = gev.load_homogeneous_template(os.path.join(ROOT, "tests", "golden", "sc1_crystal.dat"))
    specs = en.draw(args.ensemble_boxes, seed=2025)
    with en.Ensemble(specs, tpl, form=gev.FORM_FP32) as E:
        t0 = time.perf_counter()
        E.generate()
        for B, _ in E.batches.values():
            B.synchronize()
        t_ic = time.perf_counter() - t0
        print("ensemble: %d boxes generated in %.2f s (%.1f GB)" % (len(specs), t_ic, E.footprint_bytes() / 1e9),
              file=sys.stderr, flush=True)
        tuned = E.autotune()
        t_run = E.run(graph=True)
        res = [E.result(s.index) for s in specs]
    steps = sum(r["particle_steps"] for r in res)
    nbytes = sum(stage_bytes(s.ngrid, s.ngrid ** 3, 4) * r["cycles"] for s, r in zip(specs, res))
    print("ensemble: batched run %.2f s, %.3e particle-steps/s" % (t_run, steps / t_run), file=sys.stderr, flush=True)
    # the same boxes one after another, each alone; the first box of each
    # grid size also in the FP64 form, for the precision form's P(k) error
    t_alone, same = 0.0, True
    check = set(range(0, len(specs), max(1, len(specs) // 8)))
    prec = {}
    for s in specs:
        sim, st, ic, tabs = en.setup(s)
        with gev.Box(s.ngrid, s.ngrid ** 3) as b64, gev.Box(s.ngrid, s.ngrid ** 3, form=gev.FORM_FP32) as b:
            b64.generateIC_basic(sim, ic, tpl, *tabs)
            b.copy_state_from(b64)
            b.synchronize()
            t0 = time.perf_counter()
            out = b.run_to_outputs(sim, st, [0.0], s.ngrid // 2, s.boxsize)
            t_alone += time.perf_counter() - t0
            if str(s.ngrid) not in prec:
                sim64, st64, _, _ = en.setup(s)
                o64 = b64.run_to_outputs(sim64, st64, [0.0], s.ngrid // 2, s.boxsize)
                r = res[s.index]
                low = (o64["occupation"][0] > 0) & (o64["k"][0] < 0.5 * np.pi * s.ngrid / s.boxsize)
                prec[str(s.ngrid)] = {
                    "box": s.index, "cycles_fp64": st64.cycle, "cycles_fp32": r["cycles"],
                    "pk_max_rel_err_below_half_nyquist": float(np.max(np.abs(r["pk"][low] - o64["pk"][0][low]) /
                                                                      np.abs(o64["pk"][0][low])))}
        if s.index in check:
            r = res[s.index]
            same = same and st.cycle == r["cycles"] and np.array_equal(out["pk"][0], r["pk"])
    print("ensemble: boxes alone %.2f s" % t_alone, file=sys.stderr, flush=True)
    per_n = {}
    for s, r in zip(specs, res):
        d = per_n.setdefault(str(s.ngrid), {"boxes": 0, "particle_steps": 0})
        d["boxes"] += 1
        d["particle_steps"] += r["particle_steps"]
    ach = nbytes / t_run / 1e9
    return {"boxes": len(specs), "grids": [32, 48, 64, 96, 128], "form": "fp32",
            "what": "one batch per grid size, one launch per kernel over its boxes, HIP graphs per stage",
            "ic_s": round(t_ic, 3), "time_to_z0_s": round(t_run, 3), "particle_steps": int(steps),
            "value": round(steps / t_run, 1), "unit": "particle-steps/s",
            "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(ach / HBM_PEAK_GBS, 4), "algorithmic_bytes": int(nbytes)},
            "boxes_alone_s": round(t_alone, 3), "speedup_vs_alone": round(t_alone / t_run, 3),
            "identical_to_alone_sample": bool(same), "per_grid": per_n,
            "precision_vs_fp64": prec,
            "autotuned_fft_tiles": {str(N): t for N, t in tuned.items()},
            "cycles": sorted({r["cycles"] for r in res})}


def arg_parser():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=117, help="timed cycles from z = 100 (117: the whole run to z = 0)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ngrid", type=int, default=512)
    ap.add_argument("--boxsize", type=float, default=None, help="Mpc/h (default 4 * ngrid)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-timers", action="store_true", help="time without per-kernel events")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "hbm_traffic_gr_run_512_fp64.txt"),
                    help="measured HBM bytes per launch (rocprofv3 --pmc summary)")
    ap.add_argument("--state", choices=["ic", "lattice"], default="ic",
                    help="initial state: device IC generator (default) or the jittered lattice")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--no-precision-form", action="store_true",
                    help="skip the FP32 form's timing and error report")
    ap.add_argument("--no-full-run", action="store_true",
                    help="skip the whole run from z = 100 to z = 0 (both forms)")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip serving Poisson traces of whole simulations against back-to-back plain runs")
    ap.add_argument("--serving-requests", type=int, default=48)
    ap.add_argument("--serving-rates", default="12,48", help="Poisson arrival rates, requests/s")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched ensemble report")
    ap.add_argument("--no-c3", action="store_true", help="skip the 1024^3 precision-form comparison")
    ap.add_argument("--ensemble-boxes", type=int, default=256)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 5 plain 1 GiB device copies (scale for the PMC byte counters)")
    return ap


def main():
    args = arg_parser().parse_args()

    if args.calibrate_copy:
        # five launches of the library's hand-written 1 GiB copy (the PMC byte
        # counters' scale), then the probes' rates
        import gevpm as gev
        gev.bandwidth_probe(1 << 30, gev.PROBE_COPY, 5)
        print(json.dumps({"calibration": "copy", "bytes_read": 1 << 30, "bytes_written": 1 << 30,
                          "hbm_probe": hbm_probe(gev)}))
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs one box on one GPU; multi-GPU runs are not part of this build "
              "(requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import gevpm as gev
    gev.lib()
    probe = hbm_probe(gev)
    print("HBM probe: copy %.0f GB/s, read %.0f GB/s, write %.0f GB/s" %
          (probe["copy"]["GBps"], probe["read"]["GBps"], probe["write"]["GBps"]), file=sys.stderr, flush=True)

    N = args.ngrid
    Np = N ** 3
    boxsize = args.boxsize if args.boxsize else 4.0 * N
    t_setup = time.perf_counter()
    box = gev.Box(N, Np)
    ic_info = None
    if args.state == "ic":
        sim, st, ic_ms, mv0, md0 = ic_state(gev, box, N, boxsize, args.seed)
        ic_info = {"ms": round(ic_ms, 1), "particles": Np, "ngrid": N, "seed": args.seed,
                   "maxvel": mv0, "max_displacement_cells": round(md0 * N, 4),
                   "what": "generateIC_basic (ic_basic.hpp:1626-2259): 3 realisations, 15 FFTs, "
                           "T0i and Tij deposits, particle lattice, displacement and momenta"}
        print("IC generator: %.1f ms (max displacement %.3f cells, maxvel %.3e)" % (ic_ms, md0 * N, mv0),
              file=sys.stderr, flush=True)
    else:
        pos, q, ids = synthetic_state(N, args.seed)
        box.upload_particles(pos, q, ids, 0.31 / Np)
        del pos, q, ids
        sim, st = sim_setup(gev, N, boxsize)
    box.synchronize()
    t_setup = time.perf_counter() - t_setup
    print("setup %.1f s (N=%d, box %.0f Mpc/h)" % (t_setup, N, boxsize), file=sys.stderr, flush=True)

    # the CPU baseline's sample: this initial state, on the host
    cpu_state = None
    if not args.no_cpu_baseline and args.state == "ic":
        cpu_state = download_state(gev, box)

    # the FP32 precision form starts from the same state (gev_copy_state)
    box32 = st32 = None
    if not args.no_precision_form:
        box32 = gev.Box(N, Np, form=gev.FORM_FP32)

    def initial_state():
        if args.state == "ic":
            s0 = ic_state(gev, box, N, boxsize, args.seed)[1]
        else:
            p0, q0, i0 = synthetic_state(N, args.seed)
            box.upload_particles(p0, q0, i0, 0.31 / Np)
            del p0, q0, i0
            s0 = sim_setup(gev, N, boxsize)[1]
        s32 = None
        if box32 is not None:
            box32.copy_state_from(box)
            s32 = gev.CycleState(s0.a, s0.tau, s0.dtau, s0.dtau_old, s0.maxvel, s0.T00hom, s0.phi_k0, s0.cycle, 0)
        return s0, s32

    # warm-up cycles (first use of every kernel), then the initial state again
    if args.warmup > 0:
        st, st32 = initial_state()
        box.run_cycles(sim, st, args.warmup)
        if box32 is not None:
            box32.run_cycles(sim, st32, args.warmup)
        box.synchronize()
    st, st32 = initial_state()

    elapsed, kernels = timed_cycles(box, sim, st, 0, args.steps, not args.no_timers)
    print("timed %d cycles in %.3f s; a = %.5f, T00hom = %.6f, maxvel = %.3e"
          % (args.steps, elapsed, st.a, st.T00hom, st.maxvel), file=sys.stderr, flush=True)

    ms_per_step = 1e3 * elapsed / args.steps
    value = Np * args.steps / elapsed
    step_bytes = stage_bytes(N, Np)

    roofline = None
    breakdown = {}
    if kernels:
        breakdown = kernel_breakdown(kernels, args.steps, N, Np, 8)
        dom, (calls, ms) = max(kernels.items(), key=lambda kv: kv[1][1])
        avg_s = ms / calls * 1e-3
        b = algorithmic_bytes(dom, N, Np)
        traffic = measured_traffic(args.traffic, dom) if N == 512 else None
        ach = b / avg_s / 1e9
        roofline = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "algorithmic_bytes": b, "ms_per_launch": round(avg_s * 1e3, 4)}

    cpu = None
    if cpu_state is not None:
        # the box's share of the host: 16 threads (the GPU host's CPU share per GPU)
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        host = host_info(threads)
        print("cpu baseline: %d threads of %s, triad %.1f GB/s" % (threads, host["cpu_model"],
              host["stream_triad_gbs"]), file=sys.stderr, flush=True)
        v, dt, cats = cpu_baseline(cpu_state, N, boxsize, threads)
        del cpu_state
        c1 = cpu_c1_full(threads)
        print("cpu baseline: %d^3 cycle %.1f s; C1 in full %.1f s" % (N, dt, c1["main_loop_s"]),
              file=sys.stderr, flush=True)
        cpu = {"value": round(v, 1), "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "sample": "CPU oracle (FP64 C restatement of the reference cycle, OpenMP, %d threads) on the same "
                         "%d^3 workload from the device's generated initial state: cycle 1 (a full GR cycle, after "
                         "cycle 0 untimed), %.1f s" % (threads, N, dt),
               "categories_s": cats, "host": host, "c1_full_run": c1}

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": ("synthetic: device-generated Gaussian ICs (BBKS table, seed %d)" % args.seed
                 if args.state == "ic" else "synthetic: jittered lattice"),
        "config": {"workload": "GR PM cycle (main.cpp:372-879), %d^3 grid, %d^3 CDM particles, box %.0f Mpc/h, "
                               "LCDM h=0.67 Om=0.31, Cf 48, steplimit 0.04, parabolic B, FP64 plain form"
                               % (N, N, boxsize),
                   "ngrid": N, "particles": Np, "boxsize_mpc_h": boxsize, "precision": "fp64",
                   "timed_cycles": {"count": args.steps, "from_z": 100.0, "to_z": round(1.0 / st.a - 1.0, 4)}},
        "step_roofline": {"algorithmic_bytes_per_step": step_bytes,
                          "achieved": round(step_bytes / (ms_per_step * 1e-3) / 1e9, 1),
                          "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                          "min_bytes_per_step": cycle_bytes(N, Np)},
        "roofline": roofline,
        "hbm_probe": probe,
        "cpu_baseline": cpu,
        "kernels": breakdown,
        "ic_generator": ic_info,
        "precision_form": (precision_form_report(gev, box, st, box32, st32, sim, args, ms_per_step)
                           if box32 is not None else None),
    }
    if not args.no_full_run and args.state == "ic":
        out["full_run"] = full_run_report(gev, box, box32, N, boxsize, args.seed)
    if box32 is not None:
        box32.close()
    box.close()
    if not args.no_c3 and out["precision_form"] is not None:
        out["precision_form"]["c3_1024"] = c3_report(gev)
    if not args.no_ensemble:
        out["ensemble"] = ensemble_report(args)
    if not args.no_serving:
        out["serving"] = serving_report(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
