"""godunov_five_eq_amd -- MI355X-native five-equation Godunov block step.

Python view of the C ABI declared in include/godunov_five_eq.h (the product is libg5.so: HIP
kernels for gfx950 plus the host runtime).  This module only marshals arguments; every
computation runs in the library.  Loading fails loudly if libg5.so has not been built.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libg5.so")

# return codes and enums of include/godunov_five_eq.h
(G5_OK, G5_ERR_INVALID, G5_ERR_HIP, G5_ERR_UNSUPPORTED, G5_ERR_NOMEM, G5_ERR_STATE, G5_ERR_IO, G5_ERR_DIVERGED,
 G5_ERR_TOO_LARGE) = range(9)
FAIL_NONFINITE, FAIL_NEGATIVE_PRESSURE = 1, 2
BC_NONE, BC_PERIODIC, BC_ZERO_GRADIENT, BC_WALL = range(4)
TI_HANCOCK, TI_RK1, TI_RK2_SSP = range(3)
PREC_FP64, PREC_MIXED, PREC_FP64_FAST = range(3)

BC_NAMES = {"NONE": BC_NONE, "PERIODIC": BC_PERIODIC, "ZERO_GRADIENT": BC_ZERO_GRADIENT, "WALL": BC_WALL}
INLET_NONE, INLET_SHOCK_BUBBLE, INLET_JET = 0, 1, 2
SOURCE_ALLAIRE_KOKH, SOURCE_MILLER_PUCKETT = 0, 1
SOURCE_NAMES = {"ALLAIRE_KOKH": SOURCE_ALLAIRE_KOKH, "MILLER_PUCKETT": SOURCE_MILLER_PUCKETT}
INLET_NAMES = {"none": INLET_NONE, "shock_bubble": INLET_SHOCK_BUBBLE, "jet": INLET_JET}
TI_NAMES = {"HANCOCK": TI_HANCOCK, "RK1": TI_RK1, "RK2_SSP": TI_RK2_SSP}


class G5Config(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32),
        ("bsize", ctypes.c_int32),
        ("nb", ctypes.c_int32 * 3),
        ("periodic", ctypes.c_int32 * 3),
        ("bc", ctypes.c_int32 * 6),
        ("integrator", ctypes.c_int32),
        ("thinc_enabled", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("dx", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("slope_type", ctypes.c_double),
        ("smallr", ctypes.c_double),
        ("smallc", ctypes.c_double),
        ("gamma", ctypes.c_double * 2),
        ("pinf", ctypes.c_double * 2),
        ("thinc_beta", ctypes.c_double),
        ("thinc_epsilon", ctypes.c_double),
        ("xmin", ctypes.c_double * 3),
        ("inlet", ctypes.c_int32),
        ("inlet_pad", ctypes.c_int32),
        ("inlet_u", ctypes.c_double * 9),
        ("jet", ctypes.c_double * 3),
        ("source_term", ctypes.c_int32),
        ("abort_negative_pressure", ctypes.c_int32),
    ]


class G5Region(ctypes.Structure):
    _fields_ = [
        ("alpha0", ctypes.c_double),
        ("alpha_rho", ctypes.c_double * 2),
        ("p", ctypes.c_double),
        ("u", ctypes.c_double),
        ("v", ctypes.c_double),
        ("w", ctypes.c_double),
    ]


class G5Launch(ctypes.Structure):
    _fields_ = [
        ("wg_per_cu", ctypes.c_int32),
        ("xcd_remap", ctypes.c_int32),
        ("zsplit", ctypes.c_int32),
        ("ring_wave", ctypes.c_int32),
        ("col_len", ctypes.c_int32),
        ("autotuned", ctypes.c_int32),
        ("tuned_ms", ctypes.c_double),
    ]


class G5SolveSpec(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32),
        ("nb", ctypes.c_int32 * 3),
        ("kind", ctypes.c_int32),
        ("shape", ctypes.c_int32),
        ("steps", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("gamma", ctypes.c_double * 2),
        ("pinf", ctypes.c_double * 2),
        ("mach", ctypes.c_double),
        ("rho1", ctypes.c_double),
        ("p1", ctypes.c_double),
        ("rho0", ctypes.c_double),
        ("arrival", ctypes.c_double),
    ]


class G5ServeConfig(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("n_requests", ctypes.c_int32),
        ("steps_per_poll", ctypes.c_int32),
        ("rate_hz", ctypes.c_double),
        ("frac_3d", ctypes.c_double),
        ("share_3d", ctypes.c_double),
        ("hbm_budget_bytes", ctypes.c_uint64),
        ("precision", ctypes.c_int32),
        ("thinc_enabled", ctypes.c_int32),
        ("autotune", ctypes.c_int32),
        ("pad", ctypes.c_int32),
    ]


class G5ServeReport(ctypes.Structure):
    _fields_ = [
        ("completed", ctypes.c_int32),
        ("max_inflight", ctypes.c_int32),
        ("wall_s", ctypes.c_double),
        ("cell_updates", ctypes.c_double),
        ("cell_updates_per_s", ctypes.c_double),
        ("p50_ms", ctypes.c_double),
        ("p99_ms", ctypes.c_double),
        ("mean_ms", ctypes.c_double),
        ("max_ms", ctypes.c_double),
        ("last_arrival_s", ctypes.c_double),
        ("step_s_2d", ctypes.c_double),
        ("step_s_3d", ctypes.c_double),
        ("admit_s", ctypes.c_double),
        ("start_s", ctypes.c_double),
        ("wait_s", ctypes.c_double),
        ("collect_s", ctypes.c_double),
        ("failed", ctypes.c_int32),
        ("rejected", ctypes.c_int32),
    ]


SERVE_DONE, SERVE_FAILED, SERVE_REJECTED = 0, 1, 2

INDICATOR_LOHNER_SPLIT, INDICATOR_LOHNER_UNSPLIT, INDICATOR_SIMPLE_GRADIENT = 0, 1, 2
AMR_FLAG_INIT, AMR_COARSEN, AMR_NOTHING, AMR_REFINE = 0, 1, 2, 3


class G5RefineParams(ctypes.Structure):
    _fields_ = [
        ("indicator", ctypes.c_int32),
        ("level_min", ctypes.c_int32),
        ("level_max", ctypes.c_int32),
        ("base_level", ctypes.c_int32),
        ("refine_threshold", ctypes.c_double),
        ("coarsen_threshold", ctypes.c_double),
        ("epsilon_lohner", ctypes.c_double),
    ]


class G5ServeResult(ctypes.Structure):
    _fields_ = [
        ("outcome", ctypes.c_int32),
        ("failure", ctypes.c_int32),
        ("steps", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("t", ctypes.c_double),
        ("latency_ms", ctypes.c_double),
    ]


_lib = None


def lib():
    """The loaded libg5.so (raises if it is missing: there is no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        d, i32, sz = ctypes.c_double, ctypes.c_int32, ctypes.c_size_t
        vp = ctypes.c_void_p
        sig = {
            "g5_version": (ctypes.c_char_p, []),
            "g5_last_error": (ctypes.c_char_p, [vp]),
            "g5_create": (ctypes.c_int, [P(G5Config), P(vp)]),
            "g5_create_two_level": (ctypes.c_int, [P(G5Config), P(i32), i32, i32, P(vp)]),
            "g5_block_layout": (ctypes.c_int, [vp, P(i32), sz]),
            "g5_refine_flags": (ctypes.c_int, [vp, P(G5RefineParams), P(i32), P(d), sz]),
            "g5_destroy": (ctypes.c_int, [vp]),
            "g5_shape": (ctypes.c_int, [vp, P(i32), P(i32), P(i32)]),
            "g5_mem_bytes": (ctypes.c_int, [vp, P(ctypes.c_uint64)]),
            "g5_upload_state": (ctypes.c_int, [vp, P(d), sz]),
            "g5_download_state": (ctypes.c_int, [vp, P(d), sz]),
            "g5_state_device_ptr": (ctypes.c_int, [vp, P(vp)]),
            "g5_set_block_order": (ctypes.c_int, [vp, P(i32), i32]),
            "g5_do_time_step_buffers": (ctypes.c_int, [vp, vp, vp, d]),
            "g5_compute_dt_buffer": (ctypes.c_int, [vp, vp, P(d)]),
            "g5_step_noalloc": (ctypes.c_int, [P(G5Config), vp, vp, d, P(d), P(i32)]),
            "g5_compute_dt": (ctypes.c_int, [vp, P(d)]),
            "g5_fill_outside": (ctypes.c_int, [vp, P(d), sz]),
            "g5_outside_coords": (ctypes.c_int, [vp, P(i32), P(i32), sz]),
            "g5_do_time_step": (ctypes.c_int, [vp, d]),
            "g5_run": (ctypes.c_int, [vp, i32, P(d), d, P(i32)]),
            "g5_synchronize": (ctypes.c_int, [vp]),
            "g5_last_run_ms": (ctypes.c_int, [vp, P(d)]),
            "g5_time_step_kernel": (ctypes.c_int, [vp, i32, d, P(d)]),
            "g5_init_shock_bubble": (ctypes.c_int, [P(G5Config), P(G5Region), d, P(d), d, P(d), sz]),
            "g5_init_triple_point": (ctypes.c_int, [P(G5Config), P(G5Region), d, d, P(d), sz]),
            "g5_region_conservative": (ctypes.c_int, [P(G5Config), P(G5Region), P(d), sz]),
            "g5_pool_create": (ctypes.c_int, [i32, i32, i32, ctypes.c_uint64, P(vp)]),
            "g5_pool_create_rk2": (ctypes.c_int, [i32, i32, i32, ctypes.c_uint64, P(vp)]),
            "g5_pool_destroy": (ctypes.c_int, [vp]),
            "g5_pool_last_error": (ctypes.c_char_p, [vp]),
            "g5_pool_admit": (ctypes.c_int, [vp, P(G5Config), d, i32, P(i32)]),
            "g5_pool_admit_two_level": (ctypes.c_int, [vp, P(G5Config), P(i32), i32, i32, d, i32, P(i32)]),
            "g5_pool_block_layout": (ctypes.c_int, [vp, i32, P(i32), sz, P(i32)]),
            "g5_pool_release": (ctypes.c_int, [vp, i32]),
            "g5_pool_upload": (ctypes.c_int, [vp, i32, P(d), sz]),
            "g5_pool_download": (ctypes.c_int, [vp, i32, P(d), sz]),
            "g5_pool_init_shock_bubble": (ctypes.c_int, [vp, i32, P(G5Region), i32, d, P(d), P(d)]),
            "g5_pool_init_triple_point": (ctypes.c_int, [vp, i32, P(G5Region), d, d]),
            "g5_pool_start": (ctypes.c_int, [vp, P(i32), i32]),
            "g5_pool_step": (ctypes.c_int, [vp, i32]),
            "g5_pool_status": (ctypes.c_int, [vp, i32, P(d), P(i32), P(i32)]),
            "g5_pool_failure": (ctypes.c_int, [vp, i32, P(i32)]),
            "g5_pool_finished": (ctypes.c_int, [vp, P(i32), i32, P(i32)]),
            "g5_pool_usage": (ctypes.c_int, [vp, P(ctypes.c_int64), P(ctypes.c_int64), P(i32)]),
            "g5_pool_last_step_ms": (ctypes.c_int, [vp, P(d)]),
            "g5_pool_set_launch": (ctypes.c_int, [vp, ctypes.POINTER(G5Launch)]),
            "g5_pool_get_launch": (ctypes.c_int, [vp, ctypes.POINTER(G5Launch)]),
            "g5_pool_set_autotune": (ctypes.c_int, [vp, i32]),
            "g5_pool_properties": (ctypes.c_int, [vp, P(i32), P(i32), P(i32)]),
            "g5_pool_prepare": (ctypes.c_int, [vp]),
            "g5_host_alloc": (ctypes.c_int, [sz, P(vp)]),
            "g5_host_free": (ctypes.c_int, [vp]),
            "g5_pool_set_clock": (ctypes.c_int, [vp, i32, d, i32]),
            "g5_pool_step_async": (ctypes.c_int, [vp, i32]),
            "g5_batch_spec": (ctypes.c_int, [ctypes.c_uint64, i32, P(G5SolveSpec)]),
            "g5_serve_specs": (ctypes.c_int, [ctypes.c_uint64, i32, d, d, P(G5SolveSpec)]),
            "g5_spec_config": (ctypes.c_int, [P(G5SolveSpec), i32, i32, P(G5Config)]),
            "g5_spec_regions": (ctypes.c_int, [P(G5SolveSpec), P(G5Region)]),
            "g5_pool_admit_spec": (ctypes.c_int, [vp, P(G5SolveSpec), P(i32)]),
            "g5_serve": (ctypes.c_int, [P(G5ServeConfig), P(G5ServeReport), P(d), sz]),
            "g5_serve_run": (ctypes.c_int, [P(G5ServeConfig), P(G5SolveSpec), P(G5ServeReport), P(G5ServeResult),
                                            P(P(d))]),
            "g5_spec_state_size": (ctypes.c_int, [P(G5SolveSpec), P(sz)]),
            "g5_rankine_hugoniot": (ctypes.c_int, [d, d, d, d, d, d, P(d), P(d), P(d), P(d)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    """Names of the C ABI entry points (for the export check)."""
    return [
        "g5_version", "g5_last_error", "g5_create", "g5_destroy", "g5_shape", "g5_mem_bytes",
        "g5_upload_state", "g5_download_state", "g5_state_device_ptr", "g5_compute_dt",
        "g5_fill_outside", "g5_outside_coords", "g5_do_time_step", "g5_run", "g5_synchronize", "g5_last_run_ms",
        "g5_time_step_kernel", "g5_init_shock_bubble", "g5_init_triple_point", "g5_rankine_hugoniot",
        "g5_region_conservative",
        "g5_pool_create", "g5_pool_create_rk2", "g5_pool_destroy", "g5_pool_last_error", "g5_pool_admit", "g5_pool_release",
        "g5_pool_upload", "g5_pool_download", "g5_pool_init_shock_bubble", "g5_pool_init_triple_point",
        "g5_pool_start", "g5_pool_step", "g5_pool_status", "g5_pool_finished", "g5_pool_usage",
        "g5_pool_last_step_ms", "g5_pool_set_launch", "g5_pool_get_launch", "g5_pool_set_autotune",
        "g5_pool_properties", "g5_pool_prepare", "g5_host_alloc", "g5_host_free", "g5_pool_set_clock", "g5_pool_step_async", "g5_batch_spec", "g5_serve_specs", "g5_spec_config", "g5_spec_regions",
        "g5_pool_admit_spec", "g5_serve", "g5_serve_run", "g5_spec_state_size", "g5_pool_failure",
        "g5_set_block_order", "g5_do_time_step_buffers", "g5_compute_dt_buffer", "g5_step_noalloc",
        "g5_create_two_level", "g5_block_layout", "g5_refine_flags", "g5_pool_admit_two_level",
        "g5_pool_block_layout",
    ]


class G5Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"g5 error {code}: {msg}")
        self.code = code


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def nvar_of(dim):
    return 8 if dim == 2 else 9


def make_config(dim, nb, dx, *, bsize=16, periodic=(0, 0, 0), bc=("ZERO_GRADIENT",) * 6,
                integrator="HANCOCK", thinc=False, precision=PREC_FP64, cfl=0.5, slope_type=1.0,
                smallr=1e-10, smallc=1e-10, gamma=(1.4, 1.4), pinf=(0.0, 0.0), thinc_beta=1.5,
                thinc_epsilon=1e-3, xmin=(0.0, 0.0, 0.0), inlet=None, jet=(0.5, 0.5, 0.05),
                source_term="ALLAIRE_KOKH", abort_negative_pressure=False):
    """Build a G5Config; defaults follow models/HydroSettings.cpp:18-29.  inlet: None, or
    ("shock_bubble", region) / ("jet", region) -- outside cells with x < xmin (for the jet those
    whose centre is inside it, jet = (centre y, centre z, width)) take the region's state.
    source_term: "ALLAIRE_KOKH" or "MILLER_PUCKETT" (volume-fraction source term of the update).
    abort_negative_pressure: a negative pressure in a new state stops the solve (G5_ERR_DIVERGED)."""
    c = G5Config()
    c.dim = dim
    c.bsize = bsize
    nb = list(nb) + [1] * (3 - len(nb))
    for k in range(3):
        c.nb[k] = nb[k]
        c.periodic[k] = int(periodic[k]) if k < len(periodic) else 0
        c.xmin[k] = xmin[k] if k < len(xmin) else 0.0
    for f in range(6):
        b = bc[f]
        c.bc[f] = BC_NAMES[b] if isinstance(b, str) else int(b)
    c.integrator = TI_NAMES[integrator] if isinstance(integrator, str) else int(integrator)
    c.thinc_enabled = 1 if thinc else 0
    c.precision = precision
    c.dx = dx
    c.cfl = cfl
    c.slope_type = slope_type
    c.smallr = smallr
    c.smallc = smallc
    c.gamma[0], c.gamma[1] = gamma
    c.pinf[0], c.pinf[1] = pinf
    c.thinc_beta = thinc_beta
    c.thinc_epsilon = thinc_epsilon
    for k in range(3):
        c.jet[k] = jet[k]
    c.source_term = SOURCE_NAMES[source_term] if isinstance(source_term, str) else int(source_term)
    c.abort_negative_pressure = 1 if abort_negative_pressure else 0
    if inlet is not None:
        kind, reg = inlet
        c.inlet = INLET_NAMES[kind]
        u = region_conservative(c, reg)
        for k in range(u.size):
            c.inlet_u[k] = u[k]
    return c


def step_noalloc(cfg, U_dev, U2_dev, dt):
    """g5_step_noalloc on device addresses U_dev -> U2_dev; returns (next dt, failure bits)"""
    dn, f = ctypes.c_double(), ctypes.c_int32()
    rc = lib().g5_step_noalloc(ctypes.byref(cfg), ctypes.c_void_p(U_dev), ctypes.c_void_p(U2_dev), dt,
                               ctypes.byref(dn), ctypes.byref(f))
    if rc != G5_OK:
        raise G5Error(rc, "g5_step_noalloc: " + lib().g5_last_error(None).decode())
    return dn.value, f.value


def region_conservative(cfg, reg):
    """Conservative state (nvar values) of a region (common.h:68-144 get_region_init_state)."""
    u = np.zeros(nvar_of(cfg.dim), dtype=np.float64)
    rc = lib().g5_region_conservative(ctypes.byref(cfg), ctypes.byref(reg),
                                      u.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), u.size)
    if rc != 0:
        raise G5Error(rc, "g5_region_conservative")
    return u


def state_size(cfg):
    nblocks = cfg.nb[0] * cfg.nb[1] * (cfg.nb[2] if cfg.dim == 3 else 1)
    return nblocks * nvar_of(cfg.dim) * cfg.bsize ** cfg.dim


def region(alpha0, alpha_rho, p, u=0.0, v=0.0, w=0.0):
    r = G5Region()
    r.alpha0 = alpha0
    r.alpha_rho[0], r.alpha_rho[1] = alpha_rho
    r.p, r.u, r.v, r.w = p, u, v, w
    return r


def init_shock_bubble(cfg, regions, x_front, center, radius):
    """Host initial condition of InitShockBubbleDataFunctor -> state array [block][var][cell]."""
    n = state_size(cfg)
    U = np.empty(n, dtype=np.float64)
    regs = (G5Region * 3)(*regions)
    c3 = (ctypes.c_double * 3)(*(list(center) + [0.0] * (3 - len(center))))
    rc = lib().g5_init_shock_bubble(ctypes.byref(cfg), regs, x_front, c3, radius, _dptr(U), n)
    if rc != G5_OK:
        raise G5Error(rc, "g5_init_shock_bubble")
    return U


def init_triple_point(cfg, regions, xd, yd):
    """Host initial condition of InitTriplePointDataFunctor -> state array [block][var][cell]."""
    n = state_size(cfg)
    U = np.empty(n, dtype=np.float64)
    regs = (G5Region * 3)(*regions)
    rc = lib().g5_init_triple_point(ctypes.byref(cfg), regs, xd, yd, _dptr(U), n)
    if rc != G5_OK:
        raise G5Error(rc, "g5_init_triple_point")
    return U


def rankine_hugoniot(gamma, pinf, rho1, p1, mach=0.0, p2=0.0):
    """(rho2, u2, p2, shock_speed) behind a shock moving into gas at rest."""
    out = [ctypes.c_double() for _ in range(4)]
    rc = lib().g5_rankine_hugoniot(gamma, pinf, rho1, p1, mach, p2, *[ctypes.byref(o) for o in out])
    if rc != G5_OK:
        raise G5Error(rc, "g5_rankine_hugoniot")
    return tuple(o.value for o in out)


PROLONG_SIMPLE_COPY, PROLONG_EXTRAPOLATE_LINEAR_MINMOD = 0, 1
PROLONG_CONSERVATIVE_ORDER_2, PROLONG_CONSERVATIVE_ORDER_4 = 2, 3


class Context:
    """One solve on the GPU (the GodunovImplemBase surface, scheme/GodunovImplemBase.h:59-198).

    refined: (n, dim) coarse block coordinates refined into 2^dim blocks of half the cell length (a
    two-level static mesh, FP64 or FP64_FAST; see g5_create_two_level), with the given
    prolongation; None: one level."""

    def __init__(self, cfg, refined=None, prolongation=PROLONG_SIMPLE_COPY):
        self.cfg = cfg
        self._h = ctypes.c_void_p()
        if refined is None:
            rc = lib().g5_create(ctypes.byref(cfg), ctypes.byref(self._h))
        else:
            r = np.ascontiguousarray(refined, dtype=np.int32).reshape(-1, cfg.dim)
            rc = lib().g5_create_two_level(ctypes.byref(cfg), r.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                           r.shape[0], prolongation, ctypes.byref(self._h))
        if rc != G5_OK:
            raise G5Error(rc, lib().g5_last_error(None).decode())
        nb, nv, nc = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        lib().g5_shape(self._h, ctypes.byref(nb), ctypes.byref(nv), ctypes.byref(nc))
        self.nblocks, self.nvar, self.ncell = nb.value, nv.value, nc.value
        self.nstate = self.nblocks * self.nvar * self.ncell

    def _check(self, rc):
        if rc != G5_OK:
            raise G5Error(rc, lib().g5_last_error(self._h).decode())

    def close(self):
        if self._h:
            lib().g5_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64).ravel()
        self._check(lib().g5_upload_state(self._h, _dptr(U), U.size))

    def download(self):
        U = np.empty(self.nstate, dtype=np.float64)
        self._check(lib().g5_download_state(self._h, _dptr(U), U.size))
        return U

    def compute_dt(self):
        dt = ctypes.c_double()
        self._check(lib().g5_compute_dt(self._h, ctypes.byref(dt)))
        return dt.value

    def do_time_step(self, dt):
        self._check(lib().g5_do_time_step(self._h, dt))

    def set_block_order(self, coords):
        """coords: (n_blocks, 3) brick coordinates of the caller's blocks, in the caller's order"""
        c = np.ascontiguousarray(coords, dtype=np.int32).reshape(-1, 3)
        self._check(lib().g5_set_block_order(self._h, c.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), c.shape[0]))

    def refine_flags(self, indicator, level_min, level_max, base_level, refine_threshold=0.01,
                     coarsen_threshold=0.00125, epsilon_lohner=0.02, flags=None):
        """(flags, largest indicator value) per block (ComputeRefineFlags::run); flags are
        max-accumulated into the given array (AMR_FLAG_INIT when None)"""
        prm = G5RefineParams(indicator, level_min, level_max, base_level, refine_threshold, coarsen_threshold,
                             epsilon_lohner)
        f = np.zeros(self.nblocks, dtype=np.int32) if flags is None else np.array(flags, dtype=np.int32)
        imax = np.empty(self.nblocks)
        self._check(lib().g5_refine_flags(self._h, ctypes.byref(prm), f.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                          _dptr(imax), f.size))
        return f, imax

    def block_layout(self):
        """(n_blocks, 4) int32: x, y, z, level of each block of the state, in order"""
        c = np.empty((self.nblocks, 4), dtype=np.int32)
        self._check(lib().g5_block_layout(self._h, c.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), c.size))
        return c

    def device_ptr(self):
        p = ctypes.c_void_p()
        self._check(lib().g5_state_device_ptr(self._h, ctypes.byref(p)))
        return p.value

    def do_time_step_buffers(self, U_dev, U2_dev, dt):
        """one step from the caller's device array U_dev into U2_dev (integer device addresses)"""
        self._check(lib().g5_do_time_step_buffers(self._h, ctypes.c_void_p(U_dev), ctypes.c_void_p(U2_dev), dt))

    def compute_dt_buffer(self, U_dev):
        dt = ctypes.c_double()
        self._check(lib().g5_compute_dt_buffer(self._h, ctypes.c_void_p(U_dev), ctypes.byref(dt)))
        return dt.value

    def run(self, nsteps, t=0.0, t_end=1e300):
        tt = ctypes.c_double(t)
        done = ctypes.c_int32()
        self._check(lib().g5_run(self._h, nsteps, ctypes.byref(tt), t_end, ctypes.byref(done)))
        return tt.value, done.value

    def last_run_ms(self):
        ms = ctypes.c_double()
        self._check(lib().g5_last_run_ms(self._h, ctypes.byref(ms)))
        return ms.value

    def synchronize(self):
        self._check(lib().g5_synchronize(self._h))

    def mem_bytes(self):
        b = ctypes.c_uint64()
        self._check(lib().g5_mem_bytes(self._h, ctypes.byref(b)))
        return b.value

    def outside_coords(self):
        n = ctypes.c_int32()
        self._check(lib().g5_outside_coords(self._h, ctypes.byref(n), None, 0))
        coords = np.empty(3 * n.value, dtype=np.int32)
        self._check(lib().g5_outside_coords(self._h, ctypes.byref(n),
                                            coords.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                            coords.size))
        return coords.reshape(-1, 3)

    def fill_outside(self):
        n_out = self.outside_coords().shape[0]
        ring = np.empty(n_out * self.nvar * self.ncell, dtype=np.float64)
        self._check(lib().g5_fill_outside(self._h, _dptr(ring), ring.size))
        return ring

    def time_step_kernel(self, nsteps, dt):
        ms = ctypes.c_double()
        self._check(lib().g5_time_step_kernel(self._h, nsteps, dt, ctypes.byref(ms)))
        return ms.value


def batch_spec(seed, index):
    """Solve `index` of the seeded batching workload (g5_batch_spec)."""
    sp = G5SolveSpec()
    rc = lib().g5_batch_spec(seed, index, ctypes.byref(sp))
    if rc != G5_OK:
        raise G5Error(rc, "g5_batch_spec")
    return sp


def serve_specs(seed, n, rate_hz, frac_3d=0.1):
    """Requests of the seeded serving workload (g5_serve_specs)."""
    arr = (G5SolveSpec * n)()
    rc = lib().g5_serve_specs(seed, n, rate_hz, frac_3d, arr)
    if rc != G5_OK:
        raise G5Error(rc, "g5_serve_specs")
    return list(arr)


def spec_config(spec, thinc=False, precision=PREC_FP64):
    cfg = G5Config()
    rc = lib().g5_spec_config(ctypes.byref(spec), 1 if thinc else 0, precision, ctypes.byref(cfg))
    if rc != G5_OK:
        raise G5Error(rc, "g5_spec_config")
    return cfg


def spec_regions(spec):
    regs = (G5Region * 3)()
    rc = lib().g5_spec_regions(ctypes.byref(spec), regs)
    if rc != G5_OK:
        raise G5Error(rc, "g5_spec_regions")
    return list(regs)


def serve(seed, n_requests, rate_hz, *, frac_3d=0.1, share_3d=0.3, budget_bytes=200 << 30,
          precision=PREC_MIXED, thinc=False, steps_per_poll=4, autotune=False):
    """Run the serving loop (g5_serve); returns (report dict, per-request latencies in ms)."""
    c = G5ServeConfig(seed, n_requests, steps_per_poll, rate_hz, frac_3d, share_3d, int(budget_bytes),
                      precision, 1 if thinc else 0, 1 if autotune else 0, 0)
    r = G5ServeReport()
    lat = np.zeros(n_requests)
    rc = lib().g5_serve(ctypes.byref(c), ctypes.byref(r), _dptr(lat), lat.size)
    if rc != G5_OK:
        raise G5Error(rc, "g5_serve: " + lib().g5_pool_last_error(None).decode())
    rep = {k: getattr(r, k) for k, _ in G5ServeReport._fields_}
    return rep, lat


def serve_run(specs, *, budget_bytes=200 << 30, share_3d=0.3, precision=PREC_MIXED, thinc=False,
              steps_per_poll=4, autotune=False, keep_states=True):
    """Serve the given requests (g5_serve_run); returns (report dict, list of per-request result
    dicts, list of final states or None).  Arrivals must be in non-decreasing order."""
    n = len(specs)
    arr = (G5SolveSpec * n)(*specs)
    c = G5ServeConfig(0, n, steps_per_poll, 0.0, 0.0, share_3d, int(budget_bytes), precision, 1 if thinc else 0,
                      1 if autotune else 0, 0)
    r = G5ServeReport()
    res = (G5ServeResult * n)()
    states = None
    ptrs = None
    if keep_states:
        states = []
        for sp in specs:
            m = ctypes.c_size_t()
            if lib().g5_spec_state_size(ctypes.byref(sp), ctypes.byref(m)) != G5_OK:
                raise G5Error(G5_ERR_INVALID, "g5_spec_state_size")
            states.append(np.zeros(m.value))
        ptrs = (ctypes.POINTER(ctypes.c_double) * n)(*[_dptr(u) for u in states])
    rc = lib().g5_serve_run(ctypes.byref(c), arr, ctypes.byref(r), res, ptrs)
    if rc != G5_OK:
        raise G5Error(rc, "g5_serve_run: " + lib().g5_pool_last_error(None).decode())
    rep = {k: getattr(r, k) for k, _ in G5ServeReport._fields_}
    out = [{k: getattr(x, k) for k, _ in G5ServeResult._fields_ if k != "pad"} for x in res]
    return rep, out, states


class Pool:
    """Many independent solves of one dimension stepped together (g5_pool_* of the C ABI)."""

    def __init__(self, dim, thinc=False, precision=PREC_FP64, budget_bytes=8 << 30, rk2=False):
        """rk2: also hold the RK2_SSP stage buffer (g5_pool_create_rk2), for RK2_SSP solves."""
        self.dim = dim
        self._h = ctypes.c_void_p()
        create = lib().g5_pool_create_rk2 if rk2 else lib().g5_pool_create
        rc = create(dim, 1 if thinc else 0, precision, int(budget_bytes), ctypes.byref(self._h))
        if rc != G5_OK:
            raise G5Error(rc, lib().g5_pool_last_error(None).decode())
        self._cfg = {}
        self._limits = {}

    def _check(self, rc):
        if rc != G5_OK:
            raise G5Error(rc, lib().g5_pool_last_error(self._h).decode())

    def close(self):
        if self._h:
            lib().g5_pool_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def admit(self, cfg, t_end=1e300, max_steps=2 ** 31 - 1):
        sid = ctypes.c_int32()
        self._check(lib().g5_pool_admit(self._h, ctypes.byref(cfg), t_end, max_steps, ctypes.byref(sid)))
        self._cfg[sid.value] = cfg
        self._limits[sid.value] = (t_end, max_steps)
        return sid.value

    def admit_two_level(self, cfg, refined, prolongation=0, t_end=1e300, max_steps=2 ** 31 - 1):
        """admit a two-level solve (refined: (n, dim) coarse block coordinates); its state follows
        block_layout(sid)"""
        r = np.ascontiguousarray(refined, dtype=np.int32).reshape(-1, cfg.dim)
        sid = ctypes.c_int32()
        self._check(lib().g5_pool_admit_two_level(self._h, ctypes.byref(cfg),
                                                  r.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), r.shape[0],
                                                  prolongation, t_end, max_steps, ctypes.byref(sid)))
        self._cfg[sid.value] = cfg
        self._limits[sid.value] = (t_end, max_steps)
        return sid.value

    def block_layout(self, sid):
        """(n_blocks, 4) int32: x, y, z, level of each block of the solve's state"""
        n = ctypes.c_int32()
        self._check(lib().g5_pool_block_layout(self._h, sid, None, 0, ctypes.byref(n)))
        c = np.empty((n.value, 4), dtype=np.int32)
        self._check(lib().g5_pool_block_layout(self._h, sid, c.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), c.size,
                                               ctypes.byref(n)))
        return c

    def state_size(self, sid):
        n = ctypes.c_int32()
        self._check(lib().g5_pool_block_layout(self._h, sid, None, 0, ctypes.byref(n)))
        cfg = self._cfg[sid]
        return n.value * nvar_of(cfg.dim) * cfg.bsize ** cfg.dim

    def limits(self, sid):
        """(t_end, max_steps) the solve was admitted with"""
        return self._limits[sid]

    def try_admit(self, cfg, t_end=1e300, max_steps=2 ** 31 - 1):
        """admit, or None when the HBM budget has no room"""
        sid = ctypes.c_int32()
        rc = lib().g5_pool_admit(self._h, ctypes.byref(cfg), t_end, max_steps, ctypes.byref(sid))
        if rc == G5_ERR_NOMEM:
            return None
        self._check(rc)
        self._cfg[sid.value] = cfg
        self._limits[sid.value] = (t_end, max_steps)
        return sid.value

    def release(self, sid):
        self._check(lib().g5_pool_release(self._h, sid))
        self._cfg.pop(sid, None)
        self._limits.pop(sid, None)

    def upload(self, sid, U):
        U = np.ascontiguousarray(U, dtype=np.float64).ravel()
        self._check(lib().g5_pool_upload(self._h, sid, _dptr(U), U.size))

    def download(self, sid):
        U = np.empty(self.state_size(sid), dtype=np.float64)
        self._check(lib().g5_pool_download(self._h, sid, _dptr(U), U.size))
        return U

    def init_shock_bubble(self, sid, regions, x_front, center, size, shape=0):
        regs = (G5Region * 3)(*regions)
        c3 = (ctypes.c_double * 3)(*(list(center) + [0.0] * (3 - len(center))))
        sz = list(size) if hasattr(size, "__len__") else [size]
        s3 = (ctypes.c_double * 3)(*(sz + [sz[-1]] * (3 - len(sz))))
        self._check(lib().g5_pool_init_shock_bubble(self._h, sid, regs, shape, x_front, c3, s3))

    def init_triple_point(self, sid, regions, xd, yd):
        regs = (G5Region * 3)(*regions)
        self._check(lib().g5_pool_init_triple_point(self._h, sid, regs, xd, yd))

    def start(self, sids):
        arr = (ctypes.c_int32 * len(sids))(*sids)
        self._check(lib().g5_pool_start(self._h, arr, len(sids)))

    def step(self, nsteps=1):
        self._check(lib().g5_pool_step(self._h, nsteps))

    def status(self, sid):
        """(t, steps, active); raises G5Error (code G5_ERR_DIVERGED) for a failed solve"""
        t, n, a = ctypes.c_double(), ctypes.c_int32(), ctypes.c_int32()
        self._check(lib().g5_pool_status(self._h, sid, ctypes.byref(t), ctypes.byref(n), ctypes.byref(a)))
        return t.value, n.value, bool(a.value)

    def progress(self, sid):
        """(t, steps, active, G5_FAIL_* bits) of a solve, failed or not (no exception)"""
        t, n, a = ctypes.c_double(), ctypes.c_int32(), ctypes.c_int32()
        rc = lib().g5_pool_status(self._h, sid, ctypes.byref(t), ctypes.byref(n), ctypes.byref(a))
        if rc not in (G5_OK, G5_ERR_DIVERGED):
            self._check(rc)
        return t.value, n.value, bool(a.value), self.failure(sid)

    def failure(self, sid):
        """G5_FAIL_* bits of the solve (0: healthy)"""
        f = ctypes.c_int32()
        self._check(lib().g5_pool_failure(self._h, sid, ctypes.byref(f)))
        return f.value

    def finished(self):
        n = ctypes.c_int32()
        self._check(lib().g5_pool_finished(self._h, None, 0, ctypes.byref(n)))
        ids = (ctypes.c_int32 * max(1, n.value))()
        self._check(lib().g5_pool_finished(self._h, ids, n.value, ctypes.byref(n)))
        return list(ids[: n.value])

    def usage(self):
        u, c, l = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
        self._check(lib().g5_pool_usage(self._h, ctypes.byref(u), ctypes.byref(c), ctypes.byref(l)))
        return u.value, c.value, l.value

    def last_step_ms(self):
        ms = ctypes.c_double()
        self._check(lib().g5_pool_last_step_ms(self._h, ctypes.byref(ms)))
        return ms.value

    def set_launch(self, wg_per_cu=None, xcd_remap=1, zsplit=1, ring_wave=0, col_len=1):
        """Fix the step kernels' launch geometry; wg_per_cu=None returns to autotuning."""
        if wg_per_cu is None:
            self._check(lib().g5_pool_set_launch(self._h, None))
            return
        L = G5Launch(wg_per_cu, xcd_remap, zsplit, ring_wave, col_len, 0, 0.0)
        self._check(lib().g5_pool_set_launch(self._h, ctypes.byref(L)))

    def launch(self):
        L = G5Launch()
        self._check(lib().g5_pool_get_launch(self._h, ctypes.byref(L)))
        return dict(wg_per_cu=L.wg_per_cu, xcd_remap=L.xcd_remap, zsplit=L.zsplit, ring_wave=L.ring_wave,
                    col_len=L.col_len, autotuned=bool(L.autotuned), tuned_ms=L.tuned_ms)

    def set_autotune(self, enable):
        self._check(lib().g5_pool_set_autotune(self._h, 1 if enable else 0))

    def _spec_cfg(self, spec):
        d, th, pr = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        self._check(lib().g5_pool_properties(self._h, ctypes.byref(d), ctypes.byref(th), ctypes.byref(pr)))
        return spec_config(spec, bool(th.value), pr.value)

    def set_clock(self, sid, t, steps_done):
        self._check(lib().g5_pool_set_clock(self._h, sid, t, steps_done))

    def config(self, sid):
        return self._cfg[sid]

    def prepare(self):
        """Launch list, launch-geometry tuning and step-graph capture, without stepping."""
        self._check(lib().g5_pool_prepare(self._h))

    def admit_spec(self, spec):
        """Admit a workload spec and generate its initial condition on the device."""
        sid = ctypes.c_int32(-1)
        self._check(lib().g5_pool_admit_spec(self._h, ctypes.byref(spec), ctypes.byref(sid)))
        self._cfg[sid.value] = self._spec_cfg(spec)
        self._limits[sid.value] = (1e300, spec.steps)
        return sid.value

    def try_admit_spec(self, spec):
        sid = ctypes.c_int32(-1)
        rc = lib().g5_pool_admit_spec(self._h, ctypes.byref(spec), ctypes.byref(sid))
        if rc == G5_ERR_NOMEM:
            return None
        self._check(rc)
        self._cfg[sid.value] = self._spec_cfg(spec)
        self._limits[sid.value] = (1e300, spec.steps)
        return sid.value
