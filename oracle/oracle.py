"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product (the HIP library in
gevolution-1.2_amd/) never imports this module.

Array conventions are those of gev_oracle.h: real fields are numpy float64
arrays shaped (N, N, N) indexed [z, y, x]; tensor fields (6, N, N, N) in the
order 00,01,02,11,12,22; vector fields (3, N, N, N); k-space fields complex128
shaped (N, N, N//2+1) indexed [kz, ky, kx] (numpy rfftn layout); particles
pos/q float64 (n, 3), id int64 (n,).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_d = ctypes.c_double
_i = ctypes.c_int
_i64 = ctypes.c_int64
_p = ctypes.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_hconf.restype = _d
        L.orc_hconf.argtypes = [_d, _d, _p]
        L.orc_rungekutta4bg.argtypes = [_p, _d, _p, _d]
        L.orc_particle_horizon.restype = _d
        L.orc_particle_horizon.argtypes = [_d, _d, _p]
        L.orc_fft_r2c.argtypes = [_i, _p, _p]
        L.orc_fft_c2r.argtypes = [_i, _p, _p]
        L.orc_fft1d.argtypes = [_i, _p, _p, _i]
        L.orc_solve_modified_poisson_ft.argtypes = [_i, _p, _p, _d, _d]
        L.orc_project_ft_scalar.argtypes = [_i, _p, _p, _i]
        L.orc_evolve_ft_vector.argtypes = [_i, _p, _p, _d]
        L.orc_project_ft_vector.argtypes = [_i, _p, _p, _d, _d]
        L.orc_prepare_ft_source_tensor.argtypes = [_i, _p, _p, _p, _d]
        L.orc_prepare_ft_source_scalar.argtypes = [_i, _p, _p, _p, _d, _p, _d, _d, _d]
        L.orc_file_particles.argtypes = [_i, _i64, _p, _p, _p, _p, _p, _p]
        L.orc_projection_T00_project.argtypes = [_i, _i64, _p, _p, _p, _p, _d, _p, _d, _p, _d]
        L.orc_projection_Tij_project.argtypes = [_i, _i64, _p, _p, _p, _p, _d, _p, _d, _p, _d]
        L.orc_projection_T0i_project.argtypes = [_i, _i64, _p, _p, _p, _p, _d, _p, _p, _d]
        L.orc_scalar_projection_cic_project.argtypes = [_i, _i64, _p, _p, _p, _d, _p, _d]
        L.orc_update_vel.restype = _d
        L.orc_update_vel.argtypes = [_i, _i64, _p, _p, _i, _d, _p, _p, _p, _i, _p]
        L.orc_move_particles.argtypes = [_i, _i64, _p, _p, _i, _d, _p, _p, _p, _i, _p]
        L.orc_next_dtau.restype = _d
        L.orc_next_dtau.argtypes = [_p, _d]
        L.orc_box_alloc.restype = _i
        L.orc_box_alloc.argtypes = [_p, _i, _i64]
        L.orc_box_free.argtypes = [_p]
        L.orc_cycle.argtypes = [_p, _p, _p]
        L.orc_cycle_fields.argtypes = [_p, _p, _p]
        L.orc_cycle_particles.argtypes = [_p, _p, _p]
        L.orc_run_to_outputs.restype = _i
        L.orc_run_to_outputs.argtypes = [_p, _p, _p, _i, _p, _i, _d, _i, _p, _p, _p, _p, _p, _p, _p]
        L.orc_set_threads.argtypes = [_i]
        L.orc_timers_reset.argtypes = []
        L.orc_timers_get.argtypes = [_p, _p]
        L.orc_stream_triad.restype = _d
        L.orc_stream_triad.argtypes = [_i64, _i]
        _u32, _u64 = ctypes.c_uint32, ctypes.c_uint64
        L.orc_prng_block.argtypes = [_u32, _u64, _p]
        L.orc_prng_at.restype = _u32
        L.orc_prng_at.argtypes = [_u32, _u64]
        L.orc_spline_init.restype = _i
        L.orc_spline_init.argtypes = [_i, _p, _p, _p]
        L.orc_spline_eval.restype = _d
        L.orc_spline_eval.argtypes = [_i, _p, _p, _p, _d]
        L.orc_generate_cic_kernel.argtypes = [_i, _p, _i64, _p, _i]
        L.orc_generate_displacement_field.argtypes = [_i, _p, _d, _i, _p, _p, _p, _u32, _i, _i, _i, _i]
        L.orc_initialize_particle_positions.argtypes = [_i64, _p, _i, _p, _p]
        L.orc_displace_pcls_ic.restype = _d
        L.orc_displace_pcls_ic.argtypes = [_i, _i64, _p, _d, _p]
        L.orc_initialize_q_ic.restype = _d
        L.orc_initialize_q_ic.argtypes = [_i, _i64, _p, _p, _d, _p]
        L.orc_pk_primordial.restype = _d
        L.orc_pk_primordial.argtypes = [_d, _p]
        L.orc_ic_tables.argtypes = [_p, _p, _i, _p, _p, _p, _p, _p, _p, _p, _p]
        L.orc_extract_cross_spectrum.argtypes = [_i, _p, _p, _i, _i, _i, _i, _i, _i, _i, _p, _p, _p, _p, _p]
        L.orc_generate_ic_basic.restype = _i
        L.orc_generate_ic_basic.argtypes = [_p, _p, _i64, _p, _i, _p, _p, _p, _p, _p, _p, _p, _p]
        _lib = L
    return _lib


def _ptr(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays passed to the oracle must be C-contiguous"
    return a.ctypes.data


class Cosmo(ctypes.Structure):
    _fields_ = [("Omega_cdm", _d), ("Omega_b", _d), ("Omega_m", _d), ("Omega_Lambda", _d),
                ("Omega_rad", _d), ("Omega_fld", _d), ("w0_fld", _d), ("wa_fld", _d), ("h", _d)]


class Sim(ctypes.Structure):
    _fields_ = [("N", _i), ("gr_flag", _i), ("vector_flag", _i), ("Cf", _d), ("steplimit", _d),
                ("z_relax", _d), ("fourpiG", _d), ("cosmo", Cosmo)]


class State(ctypes.Structure):
    _fields_ = [("a", _d), ("tau", _d), ("dtau", _d), ("dtau_old", _d), ("maxvel", _d),
                ("T00hom", _d), ("phi_k0", _d), ("cycle", _i)]


class Box(ctypes.Structure):
    _fields_ = [("n", _i64), ("pmass", _d), ("pos", _p), ("q", _p), ("id", _p),
                ("phi", _p), ("chi", _p), ("source", _p), ("Sij", _p), ("Bi", _p),
                ("scalarFT", _p), ("SijFT", _p), ("BiFT", _p),
                ("cell", _p), ("count", _p), ("start", _p), ("perm", _p)]


class IC(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint32), ("ksphere", _i), ("correct_displacement", _i),
                ("numtile", _i), ("cr_log", _i), ("z_in", _d), ("A_s", _d), ("n_s", _d),
                ("k_pivot", _d), ("boxsize", _d)]


def make_cosmo(Omega_cdm=0.26, Omega_b=0.05, h=0.67, Omega_rad=0.0):
    om = Omega_cdm + Omega_b
    return Cosmo(Omega_cdm, Omega_b, om, 1.0 - om - Omega_rad, Omega_rad, 0.0, -1.0, 0.0, h)


# ---------------------------------------------------------------- wrappers

TIMER_NAMES = ("projections", "gravity_solver", "fft", "update_momenta", "move_particles", "spectra_output")


def timers_reset():
    lib().orc_timers_reset()


def timers():
    """{category: (seconds, calls)} since the last reset (main.cpp:907-922 categories)."""
    sec = np.zeros(len(TIMER_NAMES))
    cnt = np.zeros(len(TIMER_NAMES), dtype=np.int64)
    lib().orc_timers_get(_ptr(sec), _ptr(cnt))
    return {n: (float(s), int(c)) for n, s, c in zip(TIMER_NAMES, sec, cnt)}


def stream_triad_gbs(n=1 << 26, reps=5):
    return lib().orc_stream_triad(n, reps)


def set_threads(n):
    lib().orc_set_threads(n)


def hconf(a, fourpiG, cosmo):
    return lib().orc_hconf(a, fourpiG, ctypes.byref(cosmo))


def rk4bg(a, fourpiG, cosmo, dtau):
    x = _d(a)
    lib().orc_rungekutta4bg(ctypes.byref(x), fourpiG, ctypes.byref(cosmo), dtau)
    return x.value


def particle_horizon(a, fourpiG, cosmo):
    return lib().orc_particle_horizon(a, fourpiG, ctypes.byref(cosmo))


def fft_r2c(f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    N = f.shape[0]
    out = np.zeros((N, N, N // 2 + 1), dtype=np.complex128)
    lib().orc_fft_r2c(N, _ptr(f), _ptr(out))
    return out


def fft_c2r(k):
    k = np.ascontiguousarray(k, dtype=np.complex128)
    N = k.shape[0]
    out = np.zeros((N, N, N), dtype=np.float64)
    lib().orc_fft_c2r(N, _ptr(k), _ptr(out))
    return out


def fft1d(x, sign=-1):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros_like(x)
    lib().orc_fft1d(x.shape[0], _ptr(x), _ptr(out), sign)
    return out


def solve_modified_poisson_ft(srcFT, coeff, modif=0.0):
    srcFT = np.ascontiguousarray(srcFT, dtype=np.complex128)
    out = np.zeros_like(srcFT)
    lib().orc_solve_modified_poisson_ft(srcFT.shape[0], _ptr(srcFT), _ptr(out), coeff, modif)
    return out


def project_ft_scalar(SijFT, chiFT=None, add=0):
    SijFT = np.ascontiguousarray(SijFT, dtype=np.complex128)
    N = SijFT.shape[1]
    out = np.zeros(SijFT.shape[1:], dtype=np.complex128) if chiFT is None else np.array(chiFT, dtype=np.complex128)
    lib().orc_project_ft_scalar(N, _ptr(SijFT), _ptr(out), add)
    return out


def evolve_ft_vector(SijFT, BiFT, a2dtau):
    SijFT = np.ascontiguousarray(SijFT, dtype=np.complex128)
    out = np.array(BiFT, dtype=np.complex128, order="C")
    lib().orc_evolve_ft_vector(SijFT.shape[1], _ptr(SijFT), _ptr(out), a2dtau)
    return out


def project_ft_vector(SiFT, coeff=1.0, modif=0.0):
    SiFT = np.ascontiguousarray(SiFT, dtype=np.complex128)
    out = np.zeros_like(SiFT)
    lib().orc_project_ft_vector(SiFT.shape[1], _ptr(SiFT), _ptr(out), coeff, modif)
    return out


def prepare_ft_source_tensor(phi, Tij, coeff):
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    Tij = np.ascontiguousarray(Tij, dtype=np.float64)
    out = np.zeros_like(Tij)
    lib().orc_prepare_ft_source_tensor(phi.shape[0], _ptr(phi), _ptr(Tij), _ptr(out), coeff)
    return out


def prepare_ft_source_scalar(phi, chi, source, bgmodel, coeff, coeff2, coeff3):
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    chi = np.ascontiguousarray(chi, dtype=np.float64)
    source = np.ascontiguousarray(source, dtype=np.float64)
    out = np.zeros_like(source)
    lib().orc_prepare_ft_source_scalar(phi.shape[0], _ptr(phi), _ptr(chi), _ptr(source), bgmodel,
                                       _ptr(out), coeff, coeff2, coeff3)
    return out


def file_particles(N, pos, ids):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    ids = np.ascontiguousarray(ids, dtype=np.int64)
    n = pos.shape[0]
    cell = np.zeros(max(n, 1), dtype=np.uint32)
    count = np.zeros(N ** 3, dtype=np.uint32)
    start = np.zeros(N ** 3 + 1, dtype=np.uint64)
    perm = np.zeros(max(n, 1), dtype=np.int64)
    lib().orc_file_particles(N, n, _ptr(pos), _ptr(ids), _ptr(cell), _ptr(count), _ptr(start), _ptr(perm))
    return cell[:n], count, start, perm[:n]


def _filing(N, pos, ids):
    _, _, start, perm = file_particles(N, pos, ids)
    if perm.size == 0:
        perm = np.zeros(1, dtype=np.int64)
    return start, perm


def projection_T00_project(N, pos, q, ids, pmass, a=1.0, phi=None, coeff=1.0, out=None):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    start, perm = _filing(N, pos, ids)
    out = np.zeros((N, N, N)) if out is None else out
    phi = None if phi is None else np.ascontiguousarray(phi, dtype=np.float64)
    lib().orc_projection_T00_project(N, pos.shape[0], _ptr(pos), _ptr(q), _ptr(start), _ptr(perm),
                                     pmass, _ptr(out), a, _ptr(phi), coeff)
    return out


def projection_Tij_project(N, pos, q, ids, pmass, a=1.0, phi=None, coeff=1.0, out=None):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    start, perm = _filing(N, pos, ids)
    out = np.zeros((6, N, N, N)) if out is None else out
    phi = None if phi is None else np.ascontiguousarray(phi, dtype=np.float64)
    lib().orc_projection_Tij_project(N, pos.shape[0], _ptr(pos), _ptr(q), _ptr(start), _ptr(perm),
                                     pmass, _ptr(out), a, _ptr(phi), coeff)
    return out


def projection_T0i_project(N, pos, q, ids, pmass, phi=None, coeff=1.0, out=None):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    start, perm = _filing(N, pos, ids)
    out = np.zeros((3, N, N, N)) if out is None else out
    phi = None if phi is None else np.ascontiguousarray(phi, dtype=np.float64)
    lib().orc_projection_T0i_project(N, pos.shape[0], _ptr(pos), _ptr(q), _ptr(start), _ptr(perm),
                                     pmass, _ptr(out), _ptr(phi), coeff)
    return out


def scalar_projection_cic_project(N, pos, ids, pmass, coeff=1.0, out=None):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    start, perm = _filing(N, pos, ids)
    out = np.zeros((N, N, N)) if out is None else out
    lib().orc_scalar_projection_cic_project(N, pos.shape[0], _ptr(pos), _ptr(start), _ptr(perm),
                                            pmass, _ptr(out), coeff)
    return out


Q_GR, Q_NEWTON = 0, 1
POS_GR, POS_NEWTON = 0, 1


def update_vel(N, pos, q, method, dtau, phi, chi, Bi, nfield, params):
    """Returns (new q, sqrt(max v^2/a^2))."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    qn = np.array(q, dtype=np.float64, order="C")
    prm = np.asarray(params, dtype=np.float64)
    f = [None if x is None else np.ascontiguousarray(x, dtype=np.float64) for x in (phi, chi, Bi)]
    m = lib().orc_update_vel(N, pos.shape[0], _ptr(pos), _ptr(qn), method, dtau,
                             _ptr(f[0]), _ptr(f[1]), _ptr(f[2]), nfield, _ptr(prm))
    return qn, m


def move_particles(N, pos, q, method, dtau, phi, chi, Bi, nfield, params):
    posn = np.array(pos, dtype=np.float64, order="C")
    q = np.ascontiguousarray(q, dtype=np.float64)
    prm = np.asarray(params, dtype=np.float64)
    f = [None if x is None else np.ascontiguousarray(x, dtype=np.float64) for x in (phi, chi, Bi)]
    lib().orc_move_particles(N, posn.shape[0], _ptr(posn), _ptr(q), method, dtau,
                             _ptr(f[0]), _ptr(f[1]), _ptr(f[2]), nfield, _ptr(prm))
    return posn


# ---------------------------------------------------------------- initial conditions

def prng_block(seed, counter):
    out = np.zeros(4, dtype=np.uint64)
    lib().orc_prng_block(seed, counter, _ptr(out))
    return out


def prng_at(seed, pos):
    return lib().orc_prng_at(seed, pos)


def spline_init(x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    c = np.zeros_like(x)
    if lib().orc_spline_init(len(x), _ptr(x), _ptr(y), _ptr(c)) != 0:
        raise ValueError("spline needs at least two knots")
    return c


def spline_eval(x, y, c, xq):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    c = np.ascontiguousarray(c, dtype=np.float64)
    return np.array([lib().orc_spline_eval(len(x), _ptr(x), _ptr(y), _ptr(c), float(v))
                     for v in np.atleast_1d(xq)])


def generate_cic_kernel(N, pcldata=None, numtile=1):
    ker = np.zeros((N, N, N), dtype=np.float64)
    if pcldata is None:
        lib().orc_generate_cic_kernel(N, _ptr(ker), 0, None, 1)
    else:
        p = np.ascontiguousarray(pcldata, dtype=np.float32)
        lib().orc_generate_cic_kernel(N, _ptr(ker), len(p), _ptr(p), numtile)
    return ker


def generate_displacement_field(N, potFT, coeff, x, y, c, seed, ksphere, deconvolve_f=1,
                                ignorekernel=0, cr_log=0):
    """Returns the realisation; potFT (the transformed kernel) is not modified."""
    out = np.array(potFT, dtype=np.complex128, order="C", copy=True)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    c = np.ascontiguousarray(c, dtype=np.float64)
    lib().orc_generate_displacement_field(N, _ptr(out), coeff, len(x), _ptr(x), _ptr(y), _ptr(c),
                                          seed, ksphere, deconvolve_f, ignorekernel, cr_log)
    return out


def initialize_particle_positions(pcldata, numtile):
    p = np.ascontiguousarray(pcldata, dtype=np.float32)
    n = len(p) * numtile ** 3
    pos = np.zeros((n, 3), dtype=np.float64)
    ids = np.zeros(n, dtype=np.int64)
    lib().orc_initialize_particle_positions(len(p), _ptr(p), numtile, _ptr(pos), _ptr(ids))
    return pos, ids


def displace_pcls_ic(N, pos, coeff, field):
    """In place on pos; returns the maximum displacement."""
    field = np.ascontiguousarray(field, dtype=np.float64)
    return lib().orc_displace_pcls_ic(N, len(pos), _ptr(pos), coeff, _ptr(field))


def initialize_q_ic(N, pos, coeff, field):
    field = np.ascontiguousarray(field, dtype=np.float64)
    q = np.zeros_like(pos)
    vmax = lib().orc_initialize_q_ic(N, len(pos), _ptr(pos), _ptr(q), coeff, _ptr(field))
    return q, vmax


def pk_primordial(k, ic):
    return lib().orc_pk_primordial(k, ctypes.byref(ic))


def ic_tables(sim, ic, k, d_tot, t_tot, d_cdm, t_cdm):
    arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (k, d_tot, t_tot, d_cdm, t_cdm)]
    n = len(arrs[0])
    outs = [np.zeros(n) for _ in range(3)]
    lib().orc_ic_tables(ctypes.byref(sim), ctypes.byref(ic), n, *[_ptr(a) for a in arrs],
                        *[_ptr(o) for o in outs])
    return outs


def extract_cross_spectrum(f1, f2=None, numbins=64, deconvolve=1, ktype=1, comp1=-1, comp2=-1,
                           symmetric=None):
    """extractCrossSpectrum: f1, f2 complex k fields shaped (N, N, N//2+1) or
    (ncomp, N, N, N//2+1).  Returns (kbin, power, kscatter, pscatter, occupation)."""
    f1 = np.ascontiguousarray(f1, dtype=np.complex128)
    f2 = f1 if f2 is None else np.ascontiguousarray(f2, dtype=np.complex128)
    ncomp = 1 if f1.ndim == 3 else f1.shape[0]
    N = f1.shape[-2]
    if symmetric is None:
        symmetric = ncomp == 6
    out = [np.zeros(numbins) for _ in range(4)] + [np.zeros(numbins, dtype=np.int32)]
    lib().orc_extract_cross_spectrum(N, _ptr(f1), _ptr(f2), ncomp, int(symmetric), numbins, deconvolve, ktype,
                                     comp1, comp2, *[_ptr(o) for o in out])
    return tuple(out)


def load_template(path):
    """Gadget-2 homogeneous template as loadHomogeneousTemplate returns it
    (ic_basic.hpp:191-324): type-1 positions / BoxSize, float32 (n, 3)."""
    raw = open(path, "rb").read()
    b1 = int(np.frombuffer(raw, np.int32, 1, 0)[0])
    if b1 != 256:
        raise ValueError("template header not recognised")
    npart = np.frombuffer(raw, np.uint32, 6, 4)
    box = float(np.frombuffer(raw, np.float64, 1, 4 + 128)[0])
    off = 4 + 256 + 4 + 4 + 12 * int(npart[0])
    pos = np.frombuffer(raw, np.float32, 3 * int(npart[1]), off).reshape(-1, 3)
    return (pos / np.float32(box)).astype(np.float32)


def load_transfer_functions(path, qname, boxsize, h):
    """CLASS-format table as loadTransferFunctions reads it (ic_basic.hpp:494-717):
    columns found by their header labels; returns k*boxsize, d_<q>, t_<q>*boxsize/h."""
    kcol = dcol = tcol = None
    rows = []
    for line in open(path):
        if line.startswith("#"):
            labels = line.split(":")[1:]
            for i, lab in enumerate(labels):
                if lab.startswith("k"):
                    kcol = i
                elif lab.startswith("d_" + qname):
                    dcol = i
                elif lab.startswith("t_" + qname):
                    tcol = i
            continue
        if line.strip():
            rows.append([float(v) for v in line.split()])
    a = np.array(rows)
    return a[:, kcol] * boxsize, a[:, dcol].copy(), a[:, tcol] * boxsize / h


class OracleBox:
    """Owns an orc_box and exposes its arrays as numpy views."""

    def __init__(self, N, n):
        self.N, self.n = N, n
        self.box = Box()
        if lib().orc_box_alloc(ctypes.byref(self.box), N, n) != 0:
            raise MemoryError("orc_box_alloc failed")
        H = N // 2 + 1

        def view(addr, shape, dtype):
            count = int(np.prod(shape))
            buf = (ctypes.c_char * (count * np.dtype(dtype).itemsize)).from_address(addr)
            return np.frombuffer(buf, dtype=dtype).reshape(shape)

        b = self.box
        self.pos = view(b.pos, (n, 3), np.float64)
        self.q = view(b.q, (n, 3), np.float64)
        self.id = view(b.id, (n,), np.int64)
        self.phi = view(b.phi, (N, N, N), np.float64)
        self.chi = view(b.chi, (N, N, N), np.float64)
        self.source = view(b.source, (N, N, N), np.float64)
        self.Sij = view(b.Sij, (6, N, N, N), np.float64)
        self.Bi = view(b.Bi, (3, N, N, N), np.float64)
        self.scalarFT = view(b.scalarFT, (N, N, H), np.complex128)
        self.SijFT = view(b.SijFT, (6, N, N, H), np.complex128)
        self.BiFT = view(b.BiFT, (3, N, N, H), np.complex128)
        self.cell = view(b.cell, (n,), np.uint32)
        self.count = view(b.count, (N ** 3,), np.uint32)
        self.start = view(b.start, (N ** 3 + 1,), np.uint64)
        self.perm = view(b.perm, (n,), np.int64)

    def file(self):
        lib().orc_file_particles(self.N, self.n, self.box.pos, self.box.id, self.box.cell,
                                 self.box.count, self.box.start, self.box.perm)

    def generate_ic(self, sim, ic, pcldata, k, d_tot, t_tot, d_cdm, t_cdm):
        """generateIC_basic (transfer-function path); returns (maxvel, max_displacement)."""
        p = np.ascontiguousarray(pcldata, dtype=np.float32)
        arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (k, d_tot, t_tot, d_cdm, t_cdm)]
        mv, md = _d(0.), _d(0.)
        rc = lib().orc_generate_ic_basic(ctypes.byref(sim), ctypes.byref(ic), len(p), _ptr(p), len(arrs[0]),
                                         *[_ptr(a) for a in arrs], ctypes.byref(self.box), ctypes.byref(mv),
                                         ctypes.byref(md))
        if rc != 0:
            raise ValueError("orc_generate_ic_basic failed")
        return mv.value, md.value

    def cycle(self, sim, state):
        lib().orc_cycle(ctypes.byref(sim), ctypes.byref(self.box), ctypes.byref(state))

    def cycle_fields(self, sim, state):
        lib().orc_cycle_fields(ctypes.byref(sim), ctypes.byref(self.box), ctypes.byref(state))

    def cycle_particles(self, sim, state):
        lib().orc_cycle_particles(ctypes.byref(sim), ctypes.byref(self.box), ctypes.byref(state))

    def run_to_outputs(self, sim, state, z_pk, numbins, boxsize, max_cycles=100000):
        """The main loop with delta power-spectrum outputs at z_pk (descending),
        EXACT_OUTPUT_REDSHIFTS on; returns a dict of [len(z_pk), numbins]
        tables (k, pk, sigma_k, sigma_pk, occupation) and per-output z and
        cycle, or raises if max_cycles ran first."""
        z = np.ascontiguousarray(z_pk, dtype=np.float64)
        m = len(z)
        out = {n: np.zeros((m, numbins)) for n in ("k", "pk", "sigma_k", "sigma_pk")}
        out["occupation"] = np.zeros((m, numbins), dtype=np.int32)
        out["z"] = np.zeros(m)
        out["cycle"] = np.zeros(m, dtype=np.int32)
        rc = lib().orc_run_to_outputs(ctypes.byref(sim), ctypes.byref(self.box), ctypes.byref(state), m, _ptr(z),
                                      numbins, boxsize, max_cycles, *[_ptr(out[n]) for n in
                                      ("k", "pk", "sigma_k", "sigma_pk", "occupation", "z", "cycle")])
        if rc != m:
            raise RuntimeError("orc_run_to_outputs: %d of %d outputs written" % (max(rc, 0), m))
        return out

    def close(self):
        if self.box.pos:
            lib().orc_box_free(ctypes.byref(self.box))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
