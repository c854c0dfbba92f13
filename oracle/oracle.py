"""ctypes binding of the CPU oracle (liboracle.so) and of the reference's own
compiled point functions (oracle/_ref/libref_point.so, present only where the
reference sources were available to build it).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
benchmark's CPU-baseline leg, always as the checker or the CPU baseline,
never as the measured or shipped path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
# the restatement built with FMA contraction (a perturbation probe, see Makefile)
LIB_FMA_PATH = os.path.join(HERE, "liboracle_fma.so")
REF_POINT_PATH = os.path.join(HERE, "_ref", "libref_point.so")
REF_BIN_PATH = os.path.join(HERE, "_ref", "cbns3d_ref")

_D = ctypes.c_double
_PD = ctypes.POINTER(ctypes.c_double)
_I = ctypes.c_int


class Params(ctypes.Structure):
    """SimulationParams (include/simulation_params.h:15-53) minus the string."""

    _fields_ = [(n, _D) for n in ("CFL", "t_cur", "t_end", "Pr", "gamma", "Mach", "Re", "T_inf", "L")] + [
        (n, _I) for n in ("IM", "JM", "KM", "nstep_max", "checkpoint_freq", "num_output", "is_restart")
    ] + [(n, _D) for n in ("Mach2", "gM2", "Re_inv", "p_inf", "gam1", "gam1_inv", "Pr_inv", "C_T_inf", "C_dt_v")]


def _p(a):
    return a.ctypes.data_as(_PD)


_libs = {}


def lib(path=None):
    """The restatement's library (default), or the one at `path`."""
    path = path or LIB_PATH
    if path not in _libs:
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        L.or_params_derive.argtypes = [ctypes.POINTER(Params)]
        L.or_weno5.restype = _D
        L.or_weno5.argtypes = [_D] * 5
        L.or_rec_weno5.argtypes = [_PD] * 5 + [_PD, _PD, _D, _PD]
        L.or_calc_eigenvectors.argtypes = [_D, _PD, _PD, _D, _D, _D, _PD, _PD]
        L.or_lf_flux.argtypes = [_D, _PD, _PD, _D, _D, _D, _PD]
        L.or_create.restype = ctypes.c_void_p
        L.or_create.argtypes = [ctypes.POINTER(Params)]
        L.or_destroy.argtypes = [ctypes.c_void_p]
        L.or_dims.argtypes = [ctypes.c_void_p, ctypes.POINTER(_I)]
        L.or_init_tgv.argtypes = [ctypes.c_void_p]
        L.or_set_state.argtypes = [ctypes.c_void_p, _PD, _D, _D]
        L.or_boundary_condition.argtypes = [ctypes.c_void_p, _I]
        L.or_calc_primitive.argtypes = [ctypes.c_void_p, _I]
        L.or_init_Q_p.argtypes = [ctypes.c_void_p]
        L.or_calc_dt.restype = _D
        L.or_calc_dt.argtypes = [ctypes.c_void_p]
        L.or_rec_riemann.argtypes = [ctypes.c_void_p, _I, _I]
        L.or_calc_viscous_flux_contribution.argtypes = [ctypes.c_void_p]
        L.or_update_rk3.argtypes = [ctypes.c_void_p, _D, _I]
        L.or_step.restype = _I
        L.or_step.argtypes = [ctypes.c_void_p, _PD]
        L.or_t_cur.restype = _D
        L.or_t_cur.argtypes = [ctypes.c_void_p]
        for n in ("or_Q",):
            getattr(L, n).restype = _PD
            getattr(L, n).argtypes = [ctypes.c_void_p, _I]
        L.or_prim.restype = _PD
        L.or_prim.argtypes = [ctypes.c_void_p, _I]
        L.or_face_flux.restype = _PD
        L.or_face_flux.argtypes = [ctypes.c_void_p, _I]
        L.or_dfv.restype = _PD
        L.or_dfv.argtypes = [ctypes.c_void_p]
        L.or_get_interior.argtypes = [ctypes.c_void_p, _I, _PD]
        L.or_diagnostics.argtypes = [ctypes.c_void_p, _PD, _PD]
        L.or_set_threads.argtypes = [_I]
        _libs[path] = L
    return _libs[path]


def make_params(N, Mach=0.1, Re=1600.0, Pr=0.71, gamma=1.4, T_inf=255.556, CFL=0.5,
                t_cur=0.0, t_end=20.0, nstep_max=100, L=1.0):
    """Parameters of python/taylor_green.ini with the grid and step count set."""
    if isinstance(N, int):
        N = (N, N, N)
    p = Params(CFL=CFL, t_cur=t_cur, t_end=t_end, Pr=Pr, gamma=gamma, Mach=Mach, Re=Re,
               T_inf=T_inf, L=L, IM=N[0], JM=N[1], KM=N[2], nstep_max=nstep_max,
               checkpoint_freq=nstep_max, num_output=10, is_restart=1)
    lib().or_params_derive(ctypes.byref(p))
    return p


class Block:
    """One oracle block covering the whole periodic domain."""

    def __init__(self, params, contracted=False):
        """contracted: run on the FMA-contracted build of the restatement
        (liboracle_fma.so), a one-rounding perturbation probe."""
        self.params = params
        self._L = lib(LIB_FMA_PATH if contracted else None)
        self._h = self._L.or_create(ctypes.byref(params))
        d = (_I * 6)()
        self._L.or_dims(self._h, d)
        self.IM, self.JM, self.KM, self.IM_G, self.JM_G, self.KM_G = list(d)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.or_destroy(h)
            self._h = None

    # setup
    def init_tgv(self):
        self._L.or_init_tgv(self._h)

    def set_state(self, Q_ghosted, t_cur, mu0=0.0):
        a = np.ascontiguousarray(Q_ghosted, dtype=np.float64).reshape(-1)
        assert a.size == 5 * self.IM_G * self.JM_G * self.KM_G
        self._L.or_set_state(self._h, _p(a), t_cur, mu0)

    # stepping
    def step(self):
        dt = _D(0.0)
        fin = self._L.or_step(self._h, ctypes.byref(dt))
        return None if fin else dt.value

    def run(self, nsteps):
        return [self.step() for _ in range(nsteps)]

    @property
    def t_cur(self):
        return self._L.or_t_cur(self._h)

    # method-level access
    def boundary_condition(self, which=0):
        self._L.or_boundary_condition(self._h, which)

    def calc_primitive(self, which=0):
        self._L.or_calc_primitive(self._h, which)

    def calc_dt(self):
        return self._L.or_calc_dt(self._h)

    def rec_riemann(self, which, d):
        self._L.or_rec_riemann(self._h, which, d)

    def calc_viscous_flux_contribution(self):
        self._L.or_calc_viscous_flux_contribution(self._h)

    def update_rk3(self, dt, stage):
        self._L.or_update_rk3(self._h, dt, stage)

    # arrays (copies)
    def Q_ghosted(self, which=0):
        n = 5 * self.IM_G * self.JM_G * self.KM_G
        return np.ctypeslib.as_array(self._L.or_Q(self._h, which), shape=(n,)).copy().reshape(
            self.KM_G, self.JM_G, self.IM_G, 5)

    def Q_interior(self, which=0):
        out = np.empty((self.KM, self.JM, self.IM, 5))
        self._L.or_get_interior(self._h, which, _p(out))
        return out

    def prim(self, field):
        n = self.IM_G * self.JM_G * self.KM_G
        return np.ctypeslib.as_array(self._L.or_prim(self._h, field), shape=(n,)).copy().reshape(
            self.KM_G, self.JM_G, self.IM_G)

    def face_flux(self, d):
        shp = [(self.KM, self.JM, self.IM + 1), (self.KM, self.JM + 1, self.IM), (self.KM + 1, self.JM, self.IM)][d]
        n = 5 * shp[0] * shp[1] * shp[2]
        return np.ctypeslib.as_array(self._L.or_face_flux(self._h, d), shape=(n,)).copy().reshape(*shp, 5)

    def dfv(self):
        n = 4 * self.IM * self.JM * self.KM
        return np.ctypeslib.as_array(self._L.or_dfv(self._h), shape=(n,)).copy().reshape(self.KM, self.JM, self.IM, 4)

    def diagnostics(self):
        ke, en = _D(), _D()
        self._L.or_diagnostics(self._h, ctypes.byref(ke), ctypes.byref(en))
        return ke.value, en.value


def rounding_sensitivity(N, steps):
    """Per-field relative L-inf (each conserved field against its own
    maximum) between the restatement and its FMA-contracted build after
    `steps` Taylor-Green steps at N^3: how far the reference's own
    arithmetic moves under a one-rounding perturbation that keeps every
    operation's order."""
    a = Block(make_params(N))
    b = Block(make_params(N), contracted=True)
    for x in (a, b):
        x.init_tgv()
        x.run(steps)
    A, B = a.Q_interior(0), b.Q_interior(0)
    return [float(np.abs(A[..., e] - B[..., e]).max() / np.abs(A[..., e]).max()) for e in range(5)]


def set_threads(n):
    lib().or_set_threads(int(n))


# --- point functions ---------------------------------------------------------

def weno5(a, b, c, d, e):
    return lib().or_weno5(a, b, c, d, e)


def rec_weno5(f, R, L, gam1):
    f = np.ascontiguousarray(f, dtype=np.float64).reshape(5, 5)
    R = np.ascontiguousarray(R, dtype=np.float64).reshape(25)
    L = np.ascontiguousarray(L, dtype=np.float64).reshape(25)
    out = np.empty(5)
    rows = [np.ascontiguousarray(f[c]) for c in range(5)]
    lib().or_rec_weno5(*[_p(r) for r in rows], _p(R), _p(L), gam1, _p(out))
    return out


def calc_eigenvectors(gamma, primL, primR, n):
    pl = np.ascontiguousarray(primL, dtype=np.float64)
    pr = np.ascontiguousarray(primR, dtype=np.float64)
    R = np.empty(25)
    L = np.empty(25)
    lib().or_calc_eigenvectors(gamma, _p(pl), _p(pr), n[0], n[1], n[2], _p(R), _p(L))
    return R.reshape(5, 5), L.reshape(5, 5)


def lf_flux(gamma, primR, primL, n):
    pr = np.ascontiguousarray(primR, dtype=np.float64)
    pl = np.ascontiguousarray(primL, dtype=np.float64)
    out = np.empty(5)
    lib().or_lf_flux(gamma, _p(pr), _p(pl), n[0], n[1], n[2], _p(out))
    return out


# --- the reference's compiled point functions (built here only) --------------

_ref = None


def ref_point_lib():
    """The reference's weno.cpp / eigenvectors_roe.cpp / lf_flux.cpp compiled
    from their sources (oracle/Makefile `ref`); None where not built."""
    global _ref
    if _ref is None and os.path.exists(REF_POINT_PATH):
        L = ctypes.CDLL(REF_POINT_PATH)
        L.refpt_weno5.restype = _D
        L.refpt_weno5.argtypes = [_D] * 5
        L.refpt_rec_weno5.argtypes = [_PD, _PD, _PD, _D, _PD]
        L.refpt_calc_eigenvectors.argtypes = [_D, _PD, _PD, _PD, _PD, _PD]
        L.refpt_lf_flux.argtypes = [_D, _PD, _PD, _PD, _PD]
        _ref = L
    return _ref
