"""ctypes binding of the CPU FP64 oracle (oracle/_build/libg5oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker.  The product never imports this module.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libg5oracle.so")
# the same restatement built with the C library's tanh in THINC instead of the deterministic
# expm1-based one (G5O_LIBM_TANH): measures what the choice of tanh does to whole runs
LIB_PATHS = {"": LIB_PATH, "libm_tanh": os.path.join(HERE, "_build", "libg5oracle_libm_tanh.so")}


class OracleParams(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32),
        ("bsize", ctypes.c_int32),
        ("nb", ctypes.c_int32 * 3),
        ("periodic", ctypes.c_int32 * 3),
        ("bc", ctypes.c_int32 * 6),
        ("integrator", ctypes.c_int32),
        ("thinc_enabled", ctypes.c_int32),
        ("pad0", ctypes.c_int32),
        ("dx", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("slope_type", ctypes.c_double),
        ("smallr", ctypes.c_double),
        ("smallc", ctypes.c_double),
        ("gamma", ctypes.c_double * 2),
        ("pinf", ctypes.c_double * 2),
        ("thinc_beta", ctypes.c_double),
        ("thinc_epsilon", ctypes.c_double),
        ("inlet", ctypes.c_int32),
        ("pad1", ctypes.c_int32),
        ("inlet_u", ctypes.c_double * 9),
        ("xmin", ctypes.c_double * 3),
        ("jet", ctypes.c_double * 3),
        ("source_term", ctypes.c_int32),
        ("pad2", ctypes.c_int32),
    ]


_libs = {}


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib(variant=""):
    if variant not in _libs:
        path = LIB_PATHS[variant]
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: build the oracle (make -C oracle)")
        L = ctypes.CDLL(path)
        P = ctypes.POINTER
        d, i32, vp = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        sig = {
            "g5o_create": (vp, [P(OracleParams)]),
            "g5o_create_amr": (vp, [P(OracleParams), P(i32), i32, i32]),
            "g5o_block_levels": (None, [vp, P(i32)]),
            "g5o_cons_level": (None, [vp, i32, P(i32), P(d)]),
            "g5o_refine_flags": (None, [vp, i32, i32, i32, i32, d, d, d, P(ctypes.c_int32), P(d)]),
            "g5o_destroy": (None, [vp]),
            "g5o_num_owned": (i32, [vp]),
            "g5o_num_outside": (i32, [vp]),
            "g5o_set_state": (None, [vp, P(d)]),
            "g5o_get_state": (None, [vp, P(d)]),
            "g5o_get_state_all": (None, [vp, P(d)]),
            "g5o_block_coords": (None, [vp, P(ctypes.c_int32)]),
            "g5o_fill_outside": (None, [vp]),
            "g5o_compute_inv_dt": (d, [vp]),
            "g5o_compute_dt": (d, [vp]),
            "g5o_godunov_unsplit": (None, [vp, d]),
            "g5o_run": (i32, [vp, i32, P(d), d]),
            "g5o_set_num_threads": (None, [i32]),
            "g5o_get_max_threads": (i32, []),
            "g5o_eos_pressure": (d, [P(d), P(d), d, d, d]),
            "g5o_eos_eint": (d, [P(d), P(d), d, d, d]),
            "g5o_eos_sound_speed": (d, [P(d), P(d), d, d, d]),
            "g5o_compute_primitives": (d, [P(OracleParams), P(d), P(d)]),
            "g5o_riemann_hllc": (None, [P(OracleParams), P(d), P(d), P(d), P(d)]),
            "g5o_source_factors": (None, [P(OracleParams), P(d), P(d)]),
            "g5o_slope": (d, [d, d, d, d]),
            "g5o_tanh": (d, [d]),
            "g5o_thinc_alpha": (d, [d, d, d, d, d, i32]),
            "g5o_youngs": (None, [i32, P(d), P(d)]),
            "g5o_region_state": (None, [P(OracleParams), P(d), P(d)]),
            "g5o_box_fraction_below_plane": (d, [i32, P(d), d, P(d), d]),
            "g5o_tangent_sphere_fraction": (d, [i32, P(d), d, P(d), d]),
            "g5o_init_shock_bubble": (None, [P(OracleParams), P(d), d, P(d), d, P(d)]),
            "g5o_init_triple_point": (None, [P(OracleParams), P(d), d, d, P(d)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _libs[variant] = L
    return _libs[variant]


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def params_from(cfg):
    """OracleParams with the fields of a godunov_five_eq_amd.G5Config (or any object with them)."""
    p = OracleParams()
    p.dim, p.bsize = cfg.dim, cfg.bsize
    for k in range(3):
        p.nb[k] = cfg.nb[k] if k < cfg.dim else 1
        p.periodic[k] = cfg.periodic[k]
    for f in range(6):
        p.bc[f] = cfg.bc[f]
    p.integrator = cfg.integrator
    p.thinc_enabled = cfg.thinc_enabled
    p.dx, p.cfl, p.slope_type, p.smallr, p.smallc = cfg.dx, cfg.cfl, cfg.slope_type, cfg.smallr, cfg.smallc
    p.gamma[0], p.gamma[1] = cfg.gamma[0], cfg.gamma[1]
    p.pinf[0], p.pinf[1] = cfg.pinf[0], cfg.pinf[1]
    p.thinc_beta, p.thinc_epsilon = cfg.thinc_beta, cfg.thinc_epsilon
    p.inlet = getattr(cfg, "inlet", 0)
    for k in range(9):
        p.inlet_u[k] = cfg.inlet_u[k] if p.inlet else 0.0
    for k in range(3):
        p.xmin[k] = cfg.xmin[k]
        p.jet[k] = cfg.jet[k] if p.inlet else 0.0
    p.source_term = getattr(cfg, "source_term", 0)
    return p


class Oracle:
    def __init__(self, cfg, variant="", refined=None, prolongation=0):
        """refined: coarse block coordinates (n, 3) of a two-level mesh (each refined into 2^dim
        blocks of half the cell length); prolongation 0 SIMPLE_COPY, 1 EXTRAPOLATE_LINEAR_MINMOD"""
        self.L = lib(variant)
        self.p = params_from(cfg)
        if refined is None:
            self.h = self.L.g5o_create(ctypes.byref(self.p))
        else:
            r = np.ascontiguousarray(refined, dtype=np.int32).reshape(-1, 3)
            self.h = self.L.g5o_create_amr(ctypes.byref(self.p), r.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                           r.shape[0], prolongation)
        self.nowned = self.L.g5o_num_owned(self.h)
        self.noutside = self.L.g5o_num_outside(self.h)
        self.nvar = 8 if cfg.dim == 2 else 9
        self.ncell = cfg.bsize ** cfg.dim
        self.nstate = self.nowned * self.nvar * self.ncell

    def close(self):
        if self.h:
            self.L.g5o_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64).ravel()
        assert U.size == self.nstate
        self.L.g5o_set_state(self.h, _dp(U))

    def get_state(self):
        U = np.empty(self.nstate, dtype=np.float64)
        self.L.g5o_get_state(self.h, _dp(U))
        return U

    def get_state_all(self):
        U = np.empty((self.nowned + self.noutside) * self.nvar * self.ncell, dtype=np.float64)
        self.L.g5o_get_state_all(self.h, _dp(U))
        return U

    def cons_level(self, level, g):
        """conservative state at global cell g of a level, as the ghost fill reads it"""
        gg = (ctypes.c_int32 * 3)(*(list(g) + [0] * (3 - len(g))))
        out = np.zeros(self.nvar)
        self.L.g5o_cons_level(self.h, level, gg, _dp(out))
        return out

    def refine_flags(self, indicator, level_min, level_max, base_level, refine_th, coarsen_th, eps, flags=None):
        """(flags, indicator max) per owned block; flags max-accumulated into the given array
        (0 initial, 1 coarsen, 2 nothing, 3 refine); indicator 0 LOHNER_SPLIT, 1 LOHNER_UNSPLIT,
        2 SIMPLE_GRADIENT"""
        f = np.zeros(self.nowned, dtype=np.int32) if flags is None else np.ascontiguousarray(flags, dtype=np.int32)
        imax = np.zeros(self.nowned)
        self.L.g5o_refine_flags(self.h, indicator, level_min, level_max, base_level, refine_th, coarsen_th, eps,
                                f.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _dp(imax))
        return f, imax

    def block_levels(self):
        lv = np.zeros(self.nowned, dtype=np.int32)
        self.L.g5o_block_levels(self.h, lv.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        return lv

    def block_coords(self):
        c = np.empty(3 * (self.nowned + self.noutside), dtype=np.int32)
        self.L.g5o_block_coords(self.h, c.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        return c.reshape(-1, 3)

    def fill_outside(self):
        self.L.g5o_fill_outside(self.h)

    def compute_dt(self):
        return self.L.g5o_compute_dt(self.h)

    def compute_inv_dt(self):
        return self.L.g5o_compute_inv_dt(self.h)

    def godunov_unsplit(self, dt):
        self.L.g5o_godunov_unsplit(self.h, dt)

    def run(self, nsteps, t=0.0, t_end=1e300):
        tt = ctypes.c_double(t)
        n = self.L.g5o_run(self.h, nsteps, ctypes.byref(tt), t_end)
        return tt.value, n


def set_num_threads(n):
    lib().g5o_set_num_threads(n)


def max_threads():
    return lib().g5o_get_max_threads()


def tanh(x):
    return lib().g5o_tanh(x)


def thinc_alpha(am, ac, ap, nd, beta, side):
    """THINC face value of alpha (side +1: right face, -1: left face)."""
    return lib().g5o_thinc_alpha(am, ac, ap, nd, beta, side)


def slope(q, qp, qm, slope_type):
    return lib().g5o_slope(q, qp, qm, slope_type)


def youngs(dim, stencil):
    a = np.ascontiguousarray(stencil, dtype=np.float64).ravel()
    n = np.zeros(3)
    lib().g5o_youngs(dim, _dp(a), _dp(n))
    return n[:dim]


def eos_pressure(gamma, pinf, rho, eint, a0):
    g = (ctypes.c_double * 2)(*gamma)
    pi = (ctypes.c_double * 2)(*pinf)
    return lib().g5o_eos_pressure(g, pi, rho, eint, a0)


def eos_eint(gamma, pinf, rho, p, a0):
    g = (ctypes.c_double * 2)(*gamma)
    pi = (ctypes.c_double * 2)(*pinf)
    return lib().g5o_eos_eint(g, pi, rho, p, a0)


def eos_sound_speed(gamma, pinf, rho, p, a0):
    g = (ctypes.c_double * 2)(*gamma)
    pi = (ctypes.c_double * 2)(*pinf)
    return lib().g5o_eos_sound_speed(g, pi, rho, p, a0)


def compute_primitives(cfg, u):
    p = params_from(cfg)
    uu = np.ascontiguousarray(u, dtype=np.float64)
    q = np.zeros_like(uu)
    c = lib().g5o_compute_primitives(ctypes.byref(p), _dp(uu), _dp(q))
    return q, c


def riemann_hllc(cfg, ql, qr):
    p = params_from(cfg)
    a = np.ascontiguousarray(ql, dtype=np.float64)
    b = np.ascontiguousarray(qr, dtype=np.float64)
    flux = np.zeros_like(a)
    out3 = np.zeros(3)
    lib().g5o_riemann_hllc(ctypes.byref(p), _dp(a), _dp(b), _dp(flux), _dp(out3))
    return flux, out3[0], out3[1], out3[2]


def source_factors(cfg, u):
    """(st0, st1) of the update's volume-fraction source term for one conservative state."""
    p = params_from(cfg)
    uu = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(2)
    lib().g5o_source_factors(ctypes.byref(p), _dp(uu), _dp(out))
    return out[0], out[1]


# ---- initial conditions (common.h:68-144, init/InitShockBubble.cpp:50-225,
# init/InitTriplePoint.cpp:46-133); a region is (alpha0, (alpha_rho0, alpha_rho1), p, u, v, w) ----
def _regions(regions):
    out = []
    for r in regions:
        alpha0, (ar0, ar1), p = r[0], r[1], r[2]
        vel = list(r[3:]) + [0.0] * (3 - len(r[3:]))
        out += [alpha0, ar0, ar1, p] + vel[:3]
    return np.array(out, dtype=np.float64)


def region_state(cfg, region):
    """Conservative state (nvar values) of a region (get_region_init_state)."""
    u = np.zeros(8 if cfg.dim == 2 else 9)
    r = _regions([region])
    lib().g5o_region_state(ctypes.byref(params_from(cfg)), _dp(r), _dp(u))
    return u


def box_fraction_below_plane(n, alpha, lo, dx):
    """Fraction of the box [lo, lo + dx]^d where n.x <= alpha (exact)."""
    nn = np.ascontiguousarray(n, dtype=np.float64)
    ll = np.ascontiguousarray(lo, dtype=np.float64)
    return lib().g5o_box_fraction_below_plane(nn.size, _dp(nn), alpha, _dp(ll), dx)


def tangent_sphere_fraction(lo, dx, center, radius):
    """Volume fraction of the cell inside the sphere by the tangent plane nearest its centre."""
    ll = np.ascontiguousarray(lo, dtype=np.float64)
    cc = np.ascontiguousarray(center, dtype=np.float64)
    return lib().g5o_tangent_sphere_fraction(ll.size, _dp(ll), dx, _dp(cc), radius)


def _state_size(cfg):
    nb = cfg.nb[0] * cfg.nb[1] * (cfg.nb[2] if cfg.dim == 3 else 1)
    return nb * (8 if cfg.dim == 2 else 9) * cfg.bsize ** cfg.dim


def init_shock_bubble(cfg, regions, x_front, center, radius):
    """Shock-bubble initial state [block][var][cell] (owned blocks in brick order)."""
    U = np.empty(_state_size(cfg))
    r = _regions(regions)
    c = np.array(list(center) + [0.0] * (3 - len(center)), dtype=np.float64)
    lib().g5o_init_shock_bubble(ctypes.byref(params_from(cfg)), _dp(r), x_front, _dp(c), radius, _dp(U))
    return U


def init_triple_point(cfg, regions, xd, yd):
    """Triple-point initial state [block][var][cell] (owned blocks in brick order)."""
    U = np.empty(_state_size(cfg))
    r = _regions(regions)
    lib().g5o_init_triple_point(ctypes.byref(params_from(cfg)), _dp(r), xd, yd, _dp(U))
    return U
