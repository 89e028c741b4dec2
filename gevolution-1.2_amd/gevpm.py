"""Python binding of libgevpm.so (include/gev_pm.h) through ctypes.

`Box` mirrors the call surface that gevolution's main loop drives
(main.cpp:372-879): the method names are the reference's function names and
the arguments keep their meaning.  Host arrays are numpy; device memory is
owned by the library.  Importing this module never falls back to anything:
if the HIP library is missing or fails to load, an ImportError/OSError is
raised.

Array conventions (shared with the oracle used by the tests):
  real fields  float64 (N, N, N) indexed [z, y, x]; Sij (6, N, N, N) in the
  order 00,01,02,11,12,22; Bi (3, N, N, N);
  k fields     complex128 (N, N, N//2+1) indexed [kz, ky, kx] (numpy rfftn);
  particles    pos, q float64 (n, 3); id int64 (n,).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libgevpm.so")

_d, _i, _i64, _p, _sz = ctypes.c_double, ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_size_t

PHI, CHI, SOURCE, SIJ, BI, SCALARFT, SIJFT, BIFT = range(8)
PLAN_SOURCE, PLAN_PHI, PLAN_CHI, PLAN_SIJ, PLAN_BI = range(5)
FFT_FORWARD, FFT_BACKWARD = 0, 1
Q_GR, Q_NEWTON = 0, 1
POS_GR, POS_NEWTON = 0, 1

_NCOMP = {PHI: 1, CHI: 1, SOURCE: 1, SIJ: 6, BI: 3, SCALARFT: 1, SIJFT: 6, BIFT: 3}

EXPORTED = [
    "gev_ctx_footprint_bytes", "gev_ctx_create", "gev_ctx_destroy", "gev_last_error",
    "gev_ctx_synchronize", "gev_upload_particles", "gev_download_particles",
    "gev_particle_count", "gev_download_filing", "gev_upload_field", "gev_download_field",
    "gev_projection_init", "gev_projection_T00_project", "gev_projection_Tij_project",
    "gev_projection_T00_Tij_project", "gev_projection_T0i_project",
    "gev_scalar_projection_cic_project", "gev_field_sum", "gev_prepare_ft_source_scalar",
    "gev_prepare_ft_source_tensor", "gev_fft_execute", "gev_solve_modified_poisson_ft",
    "gev_project_ft_scalar", "gev_evolve_ft_vector", "gev_project_ft_scalar_evolve_vector",
    "gev_project_ft_vector", "gev_update_vel", "gev_move_particles", "gev_kick_drift",
    "gev_hconf", "gev_rungekutta4bg", "gev_next_dtau", "gev_run_cycles", "gev_timing_enable",
    "gev_timing_reset", "gev_timing_report", "gev_prng_fill", "gev_load_homogeneous_template",
    "gev_load_transfer_functions", "gev_generate_cic_kernel", "gev_generate_displacement_field",
    "gev_initialize_particle_positions", "gev_displace_particles_ic", "gev_initialize_q_ic",
    "gev_generate_ic_basic", "gev_extract_cross_spectrum", "gev_extract_power_spectrum",
    "gev_write_power_spectrum", "gev_hibernate", "gev_restart", "gev_run_cycles_recorded",
    "gev_write_background", "gev_copy_state", "gev_run_to_outputs", "gev_extract_power_spectrum_noalloc",
    "gev_batch_footprint_bytes", "gev_batch_create", "gev_batch_destroy", "gev_batch_last_error", "gev_batch_box",
    "gev_batch_start_box", "gev_batch_stop_box", "gev_batch_get_state", "gev_batch_run_cycles",
    "gev_batch_set_outputs", "gev_batch_step_outputs", "gev_batch_run_to_outputs", "gev_batch_get_outputs",
    "gev_batch_synchronize", "gev_batch_autotune", "gev_bandwidth_probe",
]


# forms and flags of a box (include/gev_pm.h)
FORM_FP64, FORM_FP32 = 0, 1
BOX_COMPACT = 1


class BoxDesc(ctypes.Structure):
    _fields_ = [("ngrid", ctypes.c_int32), ("form", ctypes.c_int32),
                ("npart", ctypes.c_int64), ("stream", _p), ("flags", ctypes.c_int32),
                ("reserved0", ctypes.c_int32)]


class Cosmology(ctypes.Structure):
    _fields_ = [("Omega_cdm", _d), ("Omega_b", _d), ("Omega_m", _d), ("Omega_Lambda", _d),
                ("Omega_rad", _d), ("Omega_fld", _d), ("w0_fld", _d), ("wa_fld", _d), ("h", _d)]


class SimParams(ctypes.Structure):
    _fields_ = [("gr_flag", ctypes.c_int32), ("vector_flag", ctypes.c_int32), ("Cf", _d),
                ("steplimit", _d), ("z_relax", _d), ("fourpiG", _d), ("cosmo", Cosmology)]


class ICParams(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint32), ("ksphere", ctypes.c_int32),
                ("correct_displacement", ctypes.c_int32), ("numtile", ctypes.c_int32),
                ("z_in", _d), ("A_s", _d), ("n_s", _d), ("k_pivot", _d), ("boxsize", _d)]


class BackgroundRecord(ctypes.Structure):
    _fields_ = [("cycle", ctypes.c_int32), ("reserved0", ctypes.c_int32), ("tau", _d), ("a", _d),
                ("hconf_ratio", _d), ("phi_k0", _d), ("T00hom", _d)]


class BatchDesc(ctypes.Structure):
    _fields_ = [("ngrid", ctypes.c_int32), ("form", ctypes.c_int32), ("nbox", ctypes.c_int32),
                ("gr_flag", ctypes.c_int32), ("npart", ctypes.c_int64), ("stream", ctypes.c_void_p)]


class PkOutputInfo(ctypes.Structure):
    _fields_ = [("z", _d), ("cycle", ctypes.c_int32), ("interpolated", ctypes.c_int32)]


class CycleState(ctypes.Structure):
    _fields_ = [("a", _d), ("tau", _d), ("dtau", _d), ("dtau_old", _d), ("maxvel", _d),
                ("T00hom", _d), ("phi_k0", _d), ("cycle", ctypes.c_int32),
                ("stage", ctypes.c_int32)]


_lib = None


def lib():
    """Loads libgevpm.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libgevpm.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        L.gev_ctx_footprint_bytes.restype = _i64
        L.gev_ctx_footprint_bytes.argtypes = [_p]
        L.gev_ctx_create.argtypes = [_p, ctypes.POINTER(_p)]
        L.gev_ctx_destroy.argtypes = [_p]
        L.gev_last_error.restype = ctypes.c_char_p
        L.gev_last_error.argtypes = [_p]
        L.gev_ctx_synchronize.argtypes = [_p]
        L.gev_upload_particles.argtypes = [_p, _i64, _p, _p, _p, _d]
        L.gev_download_particles.argtypes = [_p, _p, _p, _p]
        L.gev_particle_count.restype = _i64
        L.gev_particle_count.argtypes = [_p]
        L.gev_download_filing.argtypes = [_p, _p, _p]
        L.gev_upload_field.argtypes = [_p, _i, _p]
        L.gev_download_field.argtypes = [_p, _i, _p]
        L.gev_copy_state.argtypes = [_p, _p]
        L.gev_projection_init.argtypes = [_p, _i]
        L.gev_projection_T00_project.argtypes = [_p, _d, _i, _d]
        L.gev_projection_Tij_project.argtypes = [_p, _d, _i, _d]
        L.gev_projection_T00_Tij_project.argtypes = [_p, _d, _d]
        L.gev_projection_T0i_project.argtypes = [_p, _i, _d]
        L.gev_scalar_projection_cic_project.argtypes = [_p, _d]
        L.gev_field_sum.argtypes = [_p, _i, ctypes.POINTER(_d)]
        L.gev_prepare_ft_source_scalar.argtypes = [_p, _d, _d, _d, _d]
        L.gev_prepare_ft_source_tensor.argtypes = [_p, _d]
        L.gev_fft_execute.argtypes = [_p, _i, _i]
        L.gev_solve_modified_poisson_ft.argtypes = [_p, _d, _d]
        L.gev_project_ft_scalar.argtypes = [_p, _i]
        L.gev_evolve_ft_vector.argtypes = [_p, _d]
        L.gev_project_ft_scalar_evolve_vector.argtypes = [_p, _d]
        L.gev_project_ft_vector.argtypes = [_p, _d, _d]
        L.gev_update_vel.argtypes = [_p, _i, _d, _i, _p, ctypes.POINTER(_d)]
        L.gev_move_particles.argtypes = [_p, _i, _d, _i, _p]
        L.gev_kick_drift.argtypes = [_p, _i, _d, _i, _p, _i, _d, _i, _p, ctypes.POINTER(_d)]
        L.gev_hconf.restype = _d
        L.gev_hconf.argtypes = [_d, _d, _p]
        L.gev_rungekutta4bg.restype = None
        L.gev_rungekutta4bg.argtypes = [ctypes.POINTER(_d), _d, _p, _d]
        L.gev_next_dtau.restype = _d
        L.gev_next_dtau.argtypes = [_p, _i, _d]
        L.gev_run_cycles.argtypes = [_p, _p, _p, _i]
        L.gev_timing_enable.argtypes = [_p, _i]
        L.gev_timing_reset.argtypes = [_p]
        L.gev_timing_report.argtypes = [_p, ctypes.c_char_p, _sz]
        L.gev_bandwidth_probe.argtypes = [_i64, _i, _i, ctypes.POINTER(_d), ctypes.POINTER(_d)]
        L.gev_prng_fill.argtypes = [_p, ctypes.c_uint32, ctypes.c_uint64, _i64, _p]
        L.gev_load_homogeneous_template.argtypes = [ctypes.c_char_p, _p, _i64, ctypes.POINTER(_i64)]
        L.gev_load_transfer_functions.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _d, _d, _p, _p, _p,
                                                  ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
        L.gev_generate_cic_kernel.argtypes = [_p, _i, _i64, _p, ctypes.c_int32]
        L.gev_generate_displacement_field.argtypes = [_p, _d, ctypes.c_int32, _p, _p, ctypes.c_uint32,
                                                      ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
        L.gev_initialize_particle_positions.argtypes = [_p, _i64, _p, ctypes.c_int32, _d]
        L.gev_displace_particles_ic.argtypes = [_p, _d, _i, ctypes.POINTER(_d)]
        L.gev_initialize_q_ic.argtypes = [_p, _d, _i, ctypes.POINTER(_d)]
        L.gev_generate_ic_basic.argtypes = [_p, _p, _p, _i64, _p, ctypes.c_int32, _p, _p, _p, _p, _p,
                                            ctypes.POINTER(_d), ctypes.POINTER(_d)]
        L.gev_run_cycles_recorded.argtypes = [_p, _p, _p, _i, _p]
        L.gev_write_background.argtypes = [ctypes.c_char_p, _p, _i]
        L.gev_hibernate.argtypes = [_p, _p, _p, ctypes.c_char_p]
        L.gev_restart.argtypes = [_p, _p, ctypes.c_char_p, _p]
        _i32p = ctypes.POINTER(ctypes.c_int32)
        L.gev_extract_cross_spectrum.argtypes = [_p, _i, _i, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                 ctypes.c_int32, ctypes.c_int32, _p, _p, _p, _p, _p]
        L.gev_extract_power_spectrum.argtypes = [_p, _i, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                 _p, _p, _p, _p, _p]
        L.gev_write_power_spectrum.argtypes = [_p, _p, _p, _p, _p, ctypes.c_int32, _d, _d, ctypes.c_char_p,
                                               ctypes.c_char_p, _d, _d]
        L.gev_extract_power_spectrum_noalloc.argtypes = [_p, _i, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                         _p, _p, _p, _p, _p]
        _i32 = ctypes.c_int32
        L.gev_batch_footprint_bytes.restype = _i64
        L.gev_batch_footprint_bytes.argtypes = [_p]
        L.gev_batch_create.argtypes = [_p, ctypes.POINTER(_p)]
        L.gev_batch_destroy.argtypes = [_p]
        L.gev_batch_last_error.restype = ctypes.c_char_p
        L.gev_batch_last_error.argtypes = [_p]
        L.gev_batch_box.argtypes = [_p, _i32, ctypes.POINTER(_p)]
        L.gev_batch_start_box.argtypes = [_p, _i32, _p, _p, _d]
        L.gev_batch_stop_box.argtypes = [_p, _i32]
        L.gev_batch_get_state.argtypes = [_p, _i32, _p]
        L.gev_batch_run_cycles.argtypes = [_p, _i32, _i32]
        L.gev_batch_set_outputs.argtypes = [_p, _i32, _p, _i32]
        L.gev_batch_step_outputs.argtypes = [_p, _i32, _p]
        L.gev_batch_run_to_outputs.argtypes = [_p, _i32, _i32]
        L.gev_batch_get_outputs.argtypes = [_p, _i32, _p, _p, _p, _p, _p, _p]
        L.gev_batch_synchronize.argtypes = [_p]
        L.gev_batch_autotune.argtypes = [_p, ctypes.POINTER(_i32), ctypes.POINTER(_i32), _p]
        L.gev_run_to_outputs.argtypes = [_p, _p, _p, ctypes.c_int32, _p, ctypes.c_int32, _d, ctypes.c_int32,
                                         _p, _p, _p, _p, _p, _p]
        _lib = L
    return _lib


class GevError(RuntimeError):
    pass


def _ptr(a):
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data


def make_cosmology(Omega_cdm=0.26, Omega_b=0.05, h=0.67, Omega_rad=0.0):
    om = Omega_cdm + Omega_b
    return Cosmology(Omega_cdm, Omega_b, om, 1.0 - om - Omega_rad, Omega_rad, 0.0, -1.0, 0.0, h)


def hconf(a, fourpiG, cosmo):
    return lib().gev_hconf(a, fourpiG, ctypes.byref(cosmo))


def rungekutta4bg(a, fourpiG, cosmo, dtau):
    x = _d(a)
    lib().gev_rungekutta4bg(ctypes.byref(x), fourpiG, ctypes.byref(cosmo), dtau)
    return x.value


def footprint_bytes(ngrid, npart, form=FORM_FP64, compact=False):
    d = BoxDesc(ngrid, form, npart, None, BOX_COMPACT if compact else 0, 0)
    return lib().gev_ctx_footprint_bytes(ctypes.byref(d))


class Box:
    """One simulation box resident on the current GPU (a gev_ctx)."""

    @classmethod
    def _view(cls, handle, ngrid, npart, form):
        """A box of a batch (owned by the batch: close() leaves it alone)."""
        b = cls.__new__(cls)
        b.N, b.cap, b.form = int(ngrid), int(npart), int(form)
        b._c = handle
        b._owned = False
        return b

    def __init__(self, ngrid, npart, stream=None, form=FORM_FP64, compact=False):
        self._owned = True
        self.N = int(ngrid)
        self.cap = int(npart)
        self.form = int(form)
        self._c = _p()
        d = BoxDesc(self.N, self.form, self.cap, stream, BOX_COMPACT if compact else 0, 0)
        rc = lib().gev_ctx_create(ctypes.byref(d), ctypes.byref(self._c))
        if rc != 0:
            raise GevError("gev_ctx_create(N=%d, npart=%d, form=%d) failed with %d"
                           % (self.N, self.cap, self.form, rc))

    # -- plumbing
    def _chk(self, rc, what):
        if rc != 0:
            raise GevError("%s failed (%d): %s" % (what, rc, lib().gev_last_error(self._c).decode()))

    def close(self):
        if self._c and getattr(self, "_owned", True):
            lib().gev_ctx_destroy(self._c)
        self._c = _p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def synchronize(self):
        self._chk(lib().gev_ctx_synchronize(self._c), "synchronize")

    # -- data movement
    def upload_particles(self, pos, q, ids, mass):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        q = np.ascontiguousarray(q, dtype=np.float64)
        ids = np.ascontiguousarray(ids, dtype=np.int64)
        n = pos.shape[0]
        if q.shape != (n, 3) or pos.shape != (n, 3) or ids.shape != (n,):
            raise ValueError("pos/q must be (n,3) and id (n,)")
        self._chk(lib().gev_upload_particles(self._c, n, _ptr(pos), _ptr(q), _ptr(ids), mass),
                  "upload_particles")

    @property
    def n(self):
        return lib().gev_particle_count(self._c)

    def download_particles(self):
        n = self.n
        pos = np.zeros((n, 3))
        q = np.zeros((n, 3))
        ids = np.zeros(n, dtype=np.int64)
        if n:
            self._chk(lib().gev_download_particles(self._c, _ptr(pos), _ptr(q), _ptr(ids)),
                      "download_particles")
        return pos, q, ids

    def download_filing(self):
        counts = np.zeros(self.N ** 3, dtype=np.uint32)
        cells = np.zeros(max(self.n, 1), dtype=np.uint32)
        self._chk(lib().gev_download_filing(self._c, _ptr(counts), _ptr(cells)), "download_filing")
        return counts, cells[:self.n]

    def _shape(self, f):
        N, nc = self.N, _NCOMP[f]
        if f >= SCALARFT:
            shp, dt = (N, N, N // 2 + 1), np.complex128
        else:
            shp, dt = (N, N, N), np.float64
        return ((nc,) + shp if nc > 1 else shp), dt

    def upload_field(self, f, arr):
        shp, dt = self._shape(f)
        a = np.ascontiguousarray(arr, dtype=dt)
        if a.shape != shp:
            raise ValueError("field %d expects shape %s, got %s" % (f, shp, a.shape))
        self._chk(lib().gev_upload_field(self._c, f, _ptr(a)), "upload_field")

    def download_field(self, f):
        shp, dt = self._shape(f)
        a = np.zeros(shp, dtype=dt)
        self._chk(lib().gev_download_field(self._c, f, _ptr(a)), "download_field")
        return a

    def copy_state_from(self, other):
        """This box takes other's particles, fields and carried state,
        converted to this box's form (gev_copy_state)."""
        self._chk(lib().gev_copy_state(other._c, self._c), "copy_state")

    # -- the reference call surface
    def projection_init(self, f):
        self._chk(lib().gev_projection_init(self._c, f), "projection_init")

    def projection_T00_project(self, a=1.0, use_phi=True, coeff=1.0):
        self._chk(lib().gev_projection_T00_project(self._c, a, int(use_phi), coeff), "projection_T00_project")

    def projection_Tij_project(self, a=1.0, use_phi=True, coeff=1.0):
        self._chk(lib().gev_projection_Tij_project(self._c, a, int(use_phi), coeff), "projection_Tij_project")

    def projection_T00_Tij_project(self, a=1.0, coeff=1.0):
        self._chk(lib().gev_projection_T00_Tij_project(self._c, a, coeff), "projection_T00_Tij_project")

    def projection_T0i_project(self, use_phi=True, coeff=1.0):
        self._chk(lib().gev_projection_T0i_project(self._c, int(use_phi), coeff), "projection_T0i_project")

    def scalarProjectionCIC_project(self, coeff=1.0):
        self._chk(lib().gev_scalar_projection_cic_project(self._c, coeff), "scalarProjectionCIC_project")

    def field_sum(self, f):
        s = _d()
        self._chk(lib().gev_field_sum(self._c, f, ctypes.byref(s)), "field_sum")
        return s.value

    def prepareFTsource_scalar(self, bgmodel, coeff, coeff2, coeff3):
        self._chk(lib().gev_prepare_ft_source_scalar(self._c, bgmodel, coeff, coeff2, coeff3),
                  "prepareFTsource(scalar)")

    def prepareFTsource_tensor(self, coeff):
        self._chk(lib().gev_prepare_ft_source_tensor(self._c, coeff), "prepareFTsource(tensor)")

    def fft(self, plan, direction):
        self._chk(lib().gev_fft_execute(self._c, plan, direction), "fft")

    def solveModifiedPoissonFT(self, coeff, modif=0.0):
        self._chk(lib().gev_solve_modified_poisson_ft(self._c, coeff, modif), "solveModifiedPoissonFT")

    def projectFTscalar(self, add=0):
        self._chk(lib().gev_project_ft_scalar(self._c, add), "projectFTscalar")

    def evolveFTvector(self, a2dtau):
        self._chk(lib().gev_evolve_ft_vector(self._c, a2dtau), "evolveFTvector")

    def projectFTscalar_evolveFTvector(self, a2dtau):
        self._chk(lib().gev_project_ft_scalar_evolve_vector(self._c, a2dtau),
                  "projectFTscalar+evolveFTvector")

    def projectFTvector(self, coeff=1.0, modif=0.0):
        self._chk(lib().gev_project_ft_vector(self._c, coeff, modif), "projectFTvector")

    def updateVel(self, method, dtau, nfield, params, want_maxvel=True):
        prm = np.ascontiguousarray(params, dtype=np.float64)
        mv = _d()
        self._chk(lib().gev_update_vel(self._c, method, dtau, nfield, _ptr(prm),
                                       ctypes.byref(mv) if want_maxvel else None), "updateVel")
        return mv.value if want_maxvel else None

    def moveParticles(self, method, dtau, nfield, params):
        prm = np.ascontiguousarray(params, dtype=np.float64)
        self._chk(lib().gev_move_particles(self._c, method, dtau, nfield, _ptr(prm)), "moveParticles")

    def kick_drift(self, q_method, dtau_kick, nfield_kick, params_kick, pos_method, dtau_drift,
                   nfield_drift, params_drift, want_maxvel=True):
        pk = np.ascontiguousarray(params_kick, dtype=np.float64)
        pd = np.ascontiguousarray(params_drift, dtype=np.float64)
        mv = _d()
        self._chk(lib().gev_kick_drift(self._c, q_method, dtau_kick, nfield_kick, _ptr(pk),
                                       pos_method, dtau_drift, nfield_drift, _ptr(pd),
                                       ctypes.byref(mv) if want_maxvel else None), "kick_drift")
        return mv.value if want_maxvel else None

    # -- initial conditions (ic_basic.hpp)
    def prng_fill(self, seed, pos0, n):
        out = np.zeros(n, dtype=np.uint32)
        self._chk(lib().gev_prng_fill(self._c, seed, pos0, n, _ptr(out)), "prng_fill")
        return out

    def generateCICKernel(self, field=SOURCE, pcldata=None, numtile=1):
        if pcldata is None:
            self._chk(lib().gev_generate_cic_kernel(self._c, field, 0, None, 1), "cic kernel")
        else:
            p = np.ascontiguousarray(pcldata, dtype=np.float32)
            self._chk(lib().gev_generate_cic_kernel(self._c, field, len(p), _ptr(p), numtile), "cic kernel")

    def generateDisplacementField(self, coeff, x, y, seed, ksphere, deconvolve_f=1, ignorekernel=0):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        self._chk(lib().gev_generate_displacement_field(self._c, coeff, len(x), _ptr(x), _ptr(y), seed,
                                                        ksphere, deconvolve_f, ignorekernel),
                  "displacement field")

    def initializeParticlePositions(self, pcldata, numtile, mass):
        p = np.ascontiguousarray(pcldata, dtype=np.float32)
        self._chk(lib().gev_initialize_particle_positions(self._c, len(p), _ptr(p), numtile, mass),
                  "particle positions")

    def displace_pcls_ic(self, coeff, field=CHI):
        m = _d()
        self._chk(lib().gev_displace_particles_ic(self._c, coeff, field, ctypes.byref(m)), "displace")
        return m.value

    def initialize_q_ic(self, coeff, field=CHI):
        m = _d()
        self._chk(lib().gev_initialize_q_ic(self._c, coeff, field, ctypes.byref(m)), "initialize q")
        return m.value

    def generateIC_basic(self, sim, ic, pcldata, k, d_tot, t_tot, d_cdm, t_cdm):
        """Returns (maxvel, max_displacement)."""
        p = np.ascontiguousarray(pcldata, dtype=np.float32)
        arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (k, d_tot, t_tot, d_cdm, t_cdm)]
        mv, md = _d(), _d()
        self._chk(lib().gev_generate_ic_basic(self._c, ctypes.byref(sim), ctypes.byref(ic), len(p), _ptr(p),
                                              len(arrs[0]), *[_ptr(a) for a in arrs], ctypes.byref(mv),
                                              ctypes.byref(md)), "generateIC_basic")
        return mv.value, md.value

    def run_cycles_recorded(self, sim, state, ncycles):
        """run_cycles plus one background record per cycle (main.cpp:520-537)."""
        rec = (BackgroundRecord * max(ncycles, 1))()
        self._chk(lib().gev_run_cycles_recorded(self._c, ctypes.byref(sim), ctypes.byref(state), ncycles, rec),
                  "run_cycles_recorded")
        return list(rec)[:ncycles]

    # -- hibernation / restart (hibernation.hpp, ic_read.hpp)
    def hibernate(self, sim, state, path):
        self._chk(lib().gev_hibernate(self._c, ctypes.byref(sim), ctypes.byref(state), path.encode()), "hibernate")

    def restart(self, sim, path):
        """Returns the CycleState to continue with."""
        st = CycleState()
        self._chk(lib().gev_restart(self._c, ctypes.byref(sim), path.encode(), ctypes.byref(st)), "restart")
        return st

    # -- power spectra (tools.hpp)
    def extractCrossSpectrum(self, f1, f2, numbins, deconvolve=True, ktype=1, comp1=-1, comp2=-1):
        """Returns (kbin, power, kscatter, pscatter, occupation)."""
        out = [np.zeros(numbins) for _ in range(4)] + [np.zeros(numbins, dtype=np.int32)]
        self._chk(lib().gev_extract_cross_spectrum(self._c, f1, f2, numbins, int(deconvolve), ktype, comp1, comp2,
                                                   *[_ptr(o) for o in out]), "cross spectrum")
        return tuple(out)

    def extractPowerSpectrum(self, f, numbins, deconvolve=True, ktype=1):
        out = [np.zeros(numbins) for _ in range(4)] + [np.zeros(numbins, dtype=np.int32)]
        self._chk(lib().gev_extract_power_spectrum(self._c, f, numbins, int(deconvolve), ktype,
                                                   *[_ptr(o) for o in out]), "power spectrum")
        return tuple(out)

    def extractPowerSpectrum_noalloc(self, f, numbins, deconvolve=True, ktype=1):
        """extractPowerSpectrum through the entry point that allocates
        nothing (static device scratch; numbins <= 256, N <= 1024)."""
        out = [np.zeros(numbins) for _ in range(4)] + [np.zeros(numbins, dtype=np.int32)]
        self._chk(lib().gev_extract_power_spectrum_noalloc(self._c, f, numbins, int(deconvolve), ktype,
                                                           *[_ptr(o) for o in out]), "power spectrum (noalloc)")
        return tuple(out)

    def run_cycles(self, sim, state, ncycles):
        self._chk(lib().gev_run_cycles(self._c, ctypes.byref(sim), ctypes.byref(state), ncycles),
                  "run_cycles")
        return state

    def run_to_outputs(self, sim, state, z_pk, numbins, boxsize, max_cycles=100000):
        """The main loop with the delta power spectrum written at each z_pk
        (descending), EXACT_OUTPUT_REDSHIFTS on (gev_run_to_outputs); returns
        a dict of [len(z_pk), numbins] tables (k, pk, sigma_k, sigma_pk,
        occupation) and per-output z, cycle and interpolated flags."""
        z = np.ascontiguousarray(z_pk, dtype=np.float64)
        m = len(z)
        out = {n: np.zeros((m, numbins)) for n in ("k", "pk", "sigma_k", "sigma_pk")}
        out["occupation"] = np.zeros((m, numbins), dtype=np.int32)
        info = (PkOutputInfo * m)()
        self._chk(lib().gev_run_to_outputs(self._c, ctypes.byref(sim), ctypes.byref(state), m, _ptr(z), numbins,
                                           float(boxsize), int(max_cycles),
                                           *[_ptr(out[n]) for n in ("k", "pk", "sigma_k", "sigma_pk", "occupation")],
                                           info), "run_to_outputs")
        out["z"] = np.array([i.z for i in info])
        out["cycle"] = np.array([i.cycle for i in info], dtype=np.int32)
        out["interpolated"] = np.array([i.interpolated for i in info], dtype=bool)
        return out

    # -- timing
    def timing(self, on=True):
        self._chk(lib().gev_timing_enable(self._c, int(on)), "timing_enable")

    def timing_reset(self):
        self._chk(lib().gev_timing_reset(self._c), "timing_reset")

    def timing_report(self):
        buf = ctypes.create_string_buffer(1 << 16)
        self._chk(lib().gev_timing_report(self._c, buf, len(buf)), "timing_report")
        out = {}
        for line in buf.value.decode().splitlines():
            name, calls, ms = line.split()
            out[name] = (int(calls), float(ms))
        return out


PROBE_COPY, PROBE_COPY_NT, PROBE_READ, PROBE_WRITE = 0, 1, 2, 3


def bandwidth_probe(nbytes=1 << 30, kind=PROBE_COPY, reps=20):
    """(GB/s, ms per launch) of a hand-written streaming kernel over buffers
    of nbytes: the device's achievable HBM bandwidth."""
    g, m = _d(), _d()
    rc = lib().gev_bandwidth_probe(int(nbytes), int(kind), int(reps), ctypes.byref(g), ctypes.byref(m))
    if rc:
        raise GevError("bandwidth_probe failed (%d)" % rc)
    return g.value, m.value


def batch_footprint_bytes(ngrid, npart, nbox, form=FORM_FP64, gr=1):
    d = BatchDesc(ngrid, form, nbox, gr, npart, None)
    return lib().gev_batch_footprint_bytes(ctypes.byref(d))


class Batch:
    """nbox boxes of one shape stepped by the same launches (gev_batch)."""

    def __init__(self, ngrid, npart, nbox, gr=1, form=FORM_FP64, stream=None):
        self.N, self.cap, self.nbox, self.gr, self.form = int(ngrid), int(npart), int(nbox), int(gr), int(form)
        self._b = _p()
        d = BatchDesc(self.N, self.form, self.nbox, self.gr, self.cap, stream)
        rc = lib().gev_batch_create(ctypes.byref(d), ctypes.byref(self._b))
        if rc != 0:
            raise GevError("gev_batch_create(N=%d, npart=%d, nbox=%d) failed with %d" % (self.N, self.cap, self.nbox, rc))
        self.numbins = 0
        self.num_pk = 0

    def _chk(self, rc, what):
        if rc != 0:
            raise GevError("%s failed (%d): %s" % (what, rc, lib().gev_batch_last_error(self._b).decode()))

    def close(self):
        if self._b:
            lib().gev_batch_destroy(self._b)
            self._b = _p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def box(self, b):
        """Box b as a Box (for IC generation, uploads, downloads, spectra)."""
        h = _p()
        self._chk(lib().gev_batch_box(self._b, b, ctypes.byref(h)), "batch_box")
        return Box._view(h, self.N, self.cap, self.form)

    def start_box(self, b, sim, state, boxsize):
        self._chk(lib().gev_batch_start_box(self._b, b, ctypes.byref(sim), ctypes.byref(state), float(boxsize)),
                  "batch_start_box")

    def stop_box(self, b):
        self._chk(lib().gev_batch_stop_box(self._b, b), "batch_stop_box")

    def state(self, b):
        st = CycleState()
        self._chk(lib().gev_batch_get_state(self._b, b, ctypes.byref(st)), "batch_get_state")
        return st

    def run_cycles(self, ncycles, graph=True):
        self._chk(lib().gev_batch_run_cycles(self._b, ncycles, int(graph)), "batch_run_cycles")

    def set_outputs(self, z_pk, numbins):
        z = np.ascontiguousarray(z_pk, dtype=np.float64)
        self._chk(lib().gev_batch_set_outputs(self._b, len(z), _ptr(z), numbins), "batch_set_outputs")
        self.num_pk, self.numbins = len(z), int(numbins)

    def step(self, graph=True):
        """One cycle with outputs; the boxes that finished in it."""
        fin = np.zeros(self.nbox, dtype=np.int32)
        self._chk(lib().gev_batch_step_outputs(self._b, int(graph), _ptr(fin)), "batch_step_outputs")
        return np.nonzero(fin)[0].tolist()

    def run_to_outputs(self, graph=True, max_cycles=100000):
        self._chk(lib().gev_batch_run_to_outputs(self._b, int(graph), max_cycles), "batch_run_to_outputs")

    def autotune(self):
        """The tile autotuner's picks: {"y_row_bytes", "z_row_bytes", "pass_ms"}."""
        y, z = ctypes.c_int32(), ctypes.c_int32()
        ms = np.zeros((2, 3), dtype=np.float32)
        self._chk(lib().gev_batch_autotune(self._b, ctypes.byref(y), ctypes.byref(z), _ptr(ms)), "batch_autotune")
        return {"y_row_bytes": y.value, "z_row_bytes": z.value, "pass_ms": ms.tolist()}

    def synchronize(self):
        """Waits for the batch; checks the device time stepping against the host's."""
        self._chk(lib().gev_batch_synchronize(self._b), "batch_synchronize")

    def outputs(self, b):
        m, nb = self.num_pk, self.numbins
        out = {n: np.zeros((m, nb)) for n in ("k", "pk", "sigma_k", "sigma_pk")}
        out["occupation"] = np.zeros((m, nb), dtype=np.int32)
        info = (PkOutputInfo * m)()
        self._chk(lib().gev_batch_get_outputs(self._b, b, *[_ptr(out[n]) for n in
                                              ("k", "pk", "sigma_k", "sigma_pk", "occupation")], info),
                  "batch_get_outputs")
        out["z"] = np.array([i.z for i in info])
        out["cycle"] = np.array([i.cycle for i in info], dtype=np.int32)
        out["interpolated"] = np.array([i.interpolated for i in info], dtype=bool)
        return out


def next_dtau(sim, ngrid, a):
    return lib().gev_next_dtau(ctypes.byref(sim), ngrid, a)


# ---------------------------------------------------------------- initial conditions (host side)

def writePowerSpectrum(spec, rescalek, rescalep, filename, description, a, z_target=-1.0):
    """writePowerSpectrum with EXACT_OUTPUT_REDSHIFTS (tools.hpp:268-346)."""
    kbin, power, kscatter, pscatter, occ = [np.ascontiguousarray(v) for v in spec]
    occ = occ.astype(np.int32)
    rc = lib().gev_write_power_spectrum(_ptr(kbin), _ptr(power), _ptr(kscatter), _ptr(pscatter), _ptr(occ),
                                        len(kbin), rescalek, rescalep, filename.encode(), description.encode(),
                                        a, z_target)
    if rc != 0:
        raise GevError("cannot write %s (%d)" % (filename, rc))


def write_background(path, records):
    """Appends background records in the reference's format (main.cpp:531-533)."""
    arr = (BackgroundRecord * max(len(records), 1))(*records)
    if lib().gev_write_background(path.encode(), arr, len(records)) != 0:
        raise GevError("cannot write %s" % path)


def load_homogeneous_template(path):
    """loadHomogeneousTemplate: float32 (n, 3) positions in units of the template box."""
    n = _i64()
    rc = lib().gev_load_homogeneous_template(path.encode(), None, 0, ctypes.byref(n))
    if n.value <= 0:
        raise GevError("cannot read template %s (%d)" % (path, rc))
    out = np.zeros((n.value, 3), dtype=np.float32)
    rc = lib().gev_load_homogeneous_template(path.encode(), _ptr(out), n.value, ctypes.byref(n))
    if rc != 0:
        raise GevError("cannot read template %s (%d)" % (path, rc))
    return out


def load_transfer_functions(path, qname, boxsize, h):
    """loadTransferFunctions: (k * boxsize, d_<q>, t_<q> * boxsize / h)."""
    cap = 1 << 16
    k, d, t = np.zeros(cap), np.zeros(cap), np.zeros(cap)
    n = ctypes.c_int32()
    rc = lib().gev_load_transfer_functions(path.encode(), qname.encode(), boxsize, h, _ptr(k), _ptr(d),
                                           _ptr(t), cap, ctypes.byref(n))
    if rc != 0:
        raise GevError("cannot read transfer functions %s (%d)" % (path, rc))
    return k[:n.value].copy(), d[:n.value].copy(), t[:n.value].copy()
