"""ctypes binding of libcbns3d_amd.so (the C ABI declared in include/cbns3d_amd.h).

The native library is required: importing this module does not load it, but
every call does, and a missing library raises instead of falling back to
anything else.
"""
import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcbns3d_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(HERE), "include", "cbns3d_amd.h")

CBNS_FP64 = 0
CBNS_FP32_FLUX = 1
CBNS_FP64_FMA = 2
CBNS_FP64_FAST = 3
CBNS_Q = 0
CBNS_QP = 1
CBNS_ERR_NOMEM = -3
CBNS_ERR_NONFINITE = -8
CBNS_ERR_CAPACITY = -9
CBNS_NOALLOC_MAX_FACES = 8192

_D = ctypes.c_double
_PD = ctypes.POINTER(ctypes.c_double)
_I = ctypes.c_int
_PI = ctypes.POINTER(ctypes.c_int)
_VP = ctypes.c_void_p


class CbnsError(RuntimeError):
    def __init__(self, status, what):
        self.status = status
        super().__init__(f"{what}: {strerror(status)} ({status})")


class Params(ctypes.Structure):
    """cbns_params: SimulationParams (include/simulation_params.h:15-53) as POD."""

    _fields_ = [(n, _D) for n in ("CFL", "t_cur", "t_end", "Pr", "gamma", "Mach", "Re", "T_inf", "L")] + [
        (n, ctypes.c_uint32) for n in ("IM", "JM", "KM", "nstep_max", "checkpoint_freq", "num_output")
    ] + [("is_restart", ctypes.c_int32)] + [
        (n, _D) for n in ("Mach2", "gM2", "Re_inv", "p_inf", "gam1", "gam1_inv", "Pr_inv", "C_T_inf", "C_dt_v")
    ]


_SIGS = {
    "cbns_params_derive": (_I, [ctypes.POINTER(Params)]),
    "cbns_device_count": (_I, [_PI]),
    "cbns_create": (_I, [ctypes.POINTER(Params), _PI, _I, ctypes.POINTER(_VP)]),
    "cbns_destroy": (_I, [_VP]),
    "cbns_num_blocks": (_I, [_VP, _PI]),
    "cbns_block_info": (_I, [_VP, _I, _PI, _PI]),
    "cbns_memory_bytes": (_I, [_VP, ctypes.POINTER(ctypes.c_uint64)]),
    "cbns_init_tgv": (_I, [_VP]),
    "cbns_set_state": (_I, [_VP, _PD, _D, _D]),
    "cbns_get_state": (_I, [_VP, _I, _PD]),
    "cbns_get_block_state": (_I, [_VP, _I, _I, _PD]),
    "cbns_set_block_state": (_I, [_VP, _I, _I, _PD]),
    "cbns_get_block_primitive": (_I, [_VP, _I, _I, _PD]),
    "cbns_get_block_mu": (_I, [_VP, _I, _PD]),
    "cbns_set_block_mu": (_I, [_VP, _I, _PD]),
    "cbns_snapshot_begin": (_I, [_VP]),
    "cbns_snapshot_wait": (_I, [_VP, _PD]),
    "cbns_snapshot_block_state": (_I, [_VP, _I, _PD]),
    "cbns_snapshot_block_primitive": (_I, [_VP, _I, _I, _PD]),
    "cbns_boundary_condition": (_I, [_VP, _I]),
    "cbns_calc_primitive": (_I, [_VP, _I]),
    "cbns_init_Q_p": (_I, [_VP]),
    "cbns_calc_dt": (_I, [_VP, _PD]),
    "cbns_rec_riemann": (_I, [_VP, _I, _I]),
    "cbns_calc_viscous_flux_contribution": (_I, [_VP]),
    "cbns_update_rk3": (_I, [_VP, _D, _I]),
    "cbns_step": (_I, [_VP, _PD, _PI]),
    "cbns_run": (_I, [_VP, _I, _PD, _PI]),
    "cbns_get_time": (_I, [_VP, _PD]),
    "cbns_set_time": (_I, [_VP, _D]),
    "cbns_diagnostics": (_I, [_VP, _PD, _PD]),
    "cbns_spectra": (_I, [_VP, _I, _PD, _PD, _PD]),
    "cbns_sync": (_I, [_VP]),
    "cbns_face_flux_eval": (_I, [_I, _I, _D, _I, _PD, _PD, _PD]),
    "cbns_face_flux_eval_noalloc": (_I, [_I, _I, _D, _I, _PD, _PD, _PD]),
    "cbns_set_profiling": (_I, [_VP, _I]),
    "cbns_set_profiling_classes": (_I, [_VP, ctypes.c_uint]),
    "cbns_set_overlap": (_I, [_VP, _I]),
    "cbns_set_viscous_workgroups": (_I, [_VP, _I]),
    "cbns_set_band": (_I, [_VP, _I]),
    "cbns_set_face_occupancy": (_I, [_VP, _I]),
    "cbns_set_tile_shape": (_I, [_VP, _I, _I]),
    "cbns_kernel_times": (_I, [_VP, _PD, ctypes.POINTER(ctypes.c_int64)]),
    "cbns_batch_create": (_I, [ctypes.c_uint64, _I, ctypes.POINTER(_VP)]),
    "cbns_batch_destroy": (_I, [_VP]),
    "cbns_batch_bytes_needed": (_I, [ctypes.POINTER(Params), _PI, ctypes.POINTER(ctypes.c_uint64)]),
    "cbns_batch_memory": (_I, [_VP, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
    "cbns_batch_add": (_I, [_VP, ctypes.POINTER(Params), _PI, _I, _PD, _D, _D, _I, _PI]),
    "cbns_batch_remove": (_I, [_VP, _I]),
    "cbns_batch_run": (_I, [_VP, _I]),
    "cbns_batch_set_budgets": (_I, [_VP, _PI, _PI, _I]),
    "cbns_batch_sync": (_I, [_VP]),
    "cbns_batch_query": (_I, [_VP, _PI]),
    "cbns_batch_pooled": (_I, [_VP, ctypes.POINTER(ctypes.c_uint64)]),
    "cbns_batch_trim": (_I, [_VP]),
    "cbns_batch_set_workgroups": (_I, [_VP, _I]),
    "cbns_batch_set_overlap": (_I, [_VP, _I]),
    "cbns_batch_set_band": (_I, [_VP, _I]),
    "cbns_batch_status": (_I, [_VP, _I, _PI, _PI, _PD, _PI]),
    "cbns_batch_poll": (_I, [_VP, _PI, _I, _PI, _PI]),
    "cbns_batch_poll_status": (_I, [_VP, _PI, _PD, _PI, _PI, _I, _PI, _PI]),
    "cbns_batch_get_state": (_I, [_VP, _I, _PD]),
    "cbns_strerror": (ctypes.c_char_p, [_I]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"native library missing: {LIB_PATH}; build it with "
                               "`python -m cbns3d_cart_mpi_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def header_functions():
    """Names of every function the public header declares."""
    src = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(cbns_\w+)\s*\(", src, re.M)))


def strerror(status):
    return lib().cbns_strerror(status).decode()


def check(status, what):
    if status != 0:
        raise CbnsError(status, what)


def _p(a):
    return a.ctypes.data_as(_PD)


def make_params(N, Mach=0.1, Re=1600.0, Pr=0.71, gamma=1.4, T_inf=255.556, CFL=0.5, t_cur=0.0, t_end=20.0,
                nstep_max=100, L=1.0):
    """Values of python/taylor_green.ini with the grid and step count set."""
    if isinstance(N, int):
        N = (N, N, N)
    p = Params(CFL=CFL, t_cur=t_cur, t_end=t_end, Pr=Pr, gamma=gamma, Mach=Mach, Re=Re, T_inf=T_inf, L=L,
               IM=N[0], JM=N[1], KM=N[2], nstep_max=nstep_max, checkpoint_freq=nstep_max, num_output=10,
               is_restart=1)
    check(lib().cbns_params_derive(ctypes.byref(p)), "cbns_params_derive")
    return p


def device_count():
    n = _I(0)
    check(lib().cbns_device_count(ctypes.byref(n)), "cbns_device_count")
    return n.value


class Solver:
    """One context: the periodic domain split into on-device blocks."""

    def __init__(self, params, dims=(1, 1, 1), precision=CBNS_FP64):
        self.params = params
        self.N = (params.IM, params.JM, params.KM)
        d = (_I * 3)(*dims)
        h = _VP()
        check(lib().cbns_create(ctypes.byref(params), d, precision, ctypes.byref(h)), "cbns_create")
        self._h = h
        self.dims = tuple(dims)

    def close(self):
        if getattr(self, "_h", None):
            lib().cbns_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # info
    @property
    def num_blocks(self):
        n = _I()
        check(lib().cbns_num_blocks(self._h, ctypes.byref(n)), "cbns_num_blocks")
        return n.value

    def block_info(self, b):
        n, o = (_I * 3)(), (_I * 3)()
        check(lib().cbns_block_info(self._h, b, n, o), "cbns_block_info")
        return tuple(n), tuple(o)

    @property
    def memory_bytes(self):
        v = ctypes.c_uint64()
        check(lib().cbns_memory_bytes(self._h, ctypes.byref(v)), "cbns_memory_bytes")
        return v.value

    # setup
    def init_tgv(self):
        check(lib().cbns_init_tgv(self._h), "cbns_init_tgv")

    def set_state(self, Q, t_cur, mu0=0.0):
        a = np.ascontiguousarray(Q, dtype=np.float64)
        assert a.size == 5 * self.N[0] * self.N[1] * self.N[2]
        check(lib().cbns_set_state(self._h, _p(a), t_cur, mu0), "cbns_set_state")

    def get_state(self, which=CBNS_Q):
        out = np.empty((self.N[2], self.N[1], self.N[0], 5))
        check(lib().cbns_get_state(self._h, which, _p(out)), "cbns_get_state")
        return out

    def get_block_state(self, b, which=CBNS_Q):
        n, _ = self.block_info(b)
        out = np.empty((n[2] + 8, n[1] + 8, n[0] + 8, 5))
        check(lib().cbns_get_block_state(self._h, b, which, _p(out)), "cbns_get_block_state")
        return out

    def set_block_state(self, b, Qg, which=CBNS_Q):
        a = np.ascontiguousarray(Qg, dtype=np.float64)
        check(lib().cbns_set_block_state(self._h, b, which, _p(a)), "cbns_set_block_state")

    def get_block_primitive(self, b, f):
        n, _ = self.block_info(b)
        out = np.empty((n[2] + 8, n[1] + 8, n[0] + 8))
        check(lib().cbns_get_block_primitive(self._h, b, f, _p(out)), "cbns_get_block_primitive")
        return out

    def get_block_mu(self, b):
        """Interior mu of block b, [k][j][i]: the viscosity the next dt reads."""
        n, _ = self.block_info(b)
        out = np.empty((n[2], n[1], n[0]))
        check(lib().cbns_get_block_mu(self._h, b, _p(out)), "cbns_get_block_mu")
        return out

    def set_block_mu(self, b, mu):
        a = np.ascontiguousarray(mu, dtype=np.float64)
        n, _ = self.block_info(b)
        assert a.size == n[0] * n[1] * n[2]
        check(lib().cbns_set_block_mu(self._h, b, _p(a)), "cbns_set_block_mu")

    def snapshot_begin(self):
        """Asynchronous snapshot of Q and the primitives (read with the
        snapshot_* methods; valid until the next snapshot_begin)."""
        check(lib().cbns_snapshot_begin(self._h), "cbns_snapshot_begin")

    def snapshot_wait(self):
        t = _D()
        check(lib().cbns_snapshot_wait(self._h, ctypes.byref(t)), "cbns_snapshot_wait")
        return t.value

    def snapshot_block_state(self, b):
        n, _ = self.block_info(b)
        out = np.empty((n[2] + 8, n[1] + 8, n[0] + 8, 5))
        check(lib().cbns_snapshot_block_state(self._h, b, _p(out)), "cbns_snapshot_block_state")
        return out

    def snapshot_block_primitive(self, b, f):
        n, _ = self.block_info(b)
        out = np.empty((n[2] + 8, n[1] + 8, n[0] + 8))
        check(lib().cbns_snapshot_block_primitive(self._h, b, f, _p(out)), "cbns_snapshot_block_primitive")
        return out

    # method level
    def boundary_condition(self, which=CBNS_Q):
        check(lib().cbns_boundary_condition(self._h, which), "cbns_boundary_condition")

    def calc_primitive(self, which=CBNS_Q):
        check(lib().cbns_calc_primitive(self._h, which), "cbns_calc_primitive")

    def init_Q_p(self):
        check(lib().cbns_init_Q_p(self._h), "cbns_init_Q_p")

    def calc_dt(self):
        dt = _D()
        check(lib().cbns_calc_dt(self._h, ctypes.byref(dt)), "cbns_calc_dt")
        return dt.value

    def rec_riemann(self, which, d):
        check(lib().cbns_rec_riemann(self._h, which, d), "cbns_rec_riemann")

    def calc_viscous_flux_contribution(self):
        check(lib().cbns_calc_viscous_flux_contribution(self._h), "cbns_calc_viscous_flux_contribution")

    def update_rk3(self, dt, stage):
        check(lib().cbns_update_rk3(self._h, dt, stage), "cbns_update_rk3")

    # fused
    def step(self):
        dt, fin = _D(), _I()
        check(lib().cbns_step(self._h, ctypes.byref(dt), ctypes.byref(fin)), "cbns_step")
        return None if fin.value else dt.value

    def run(self, nsteps):
        dts = np.zeros(max(nsteps, 1))
        done = _I()
        check(lib().cbns_run(self._h, nsteps, _p(dts), ctypes.byref(done)), "cbns_run")
        return dts[:done.value]

    @property
    def t_cur(self):
        t = _D()
        check(lib().cbns_get_time(self._h, ctypes.byref(t)), "cbns_get_time")
        return t.value

    def diagnostics(self):
        ke, en = _D(), _D()
        check(lib().cbns_diagnostics(self._h, ctypes.byref(ke), ctypes.byref(en)), "cbns_diagnostics")
        return ke.value, en.value

    def spectra(self, nshell=None):
        """Device shell spectra of the velocity of Q: (ke, enstrophy,
        dilatation), each of nshell shells (default: up to the largest
        |k| of the grid)."""
        if nshell is None:
            nshell = int(np.ceil(np.sqrt(sum((n // 2) ** 2 for n in self.N)))) + 1
        out = [np.zeros(nshell) for _ in range(3)]
        check(lib().cbns_spectra(self._h, int(nshell), *(_p(o) for o in out)), "cbns_spectra")
        return tuple(out)

    def sync(self):
        check(lib().cbns_sync(self._h), "cbns_sync")

    KERNEL_CLASSES = ("face_x", "face_y", "face_z", "viscous", "update", "ghost_fill", "primitive", "dt")

    def set_band(self, rows):
        check(lib().cbns_set_band(self._h, int(rows)), "cbns_set_band")

    def set_face_occupancy(self, waves):
        check(lib().cbns_set_face_occupancy(self._h, int(waves)), "cbns_set_face_occupancy")

    def set_tile_shape(self, visc_planes, run_cells):
        """Most planes per viscous z-chunk and cells per run of the fused z
        pass and the tolerance-bound x+y march (0: the default, 32; the x+y
        march takes at most 64)."""
        check(lib().cbns_set_tile_shape(self._h, int(visc_planes), int(run_cells)), "cbns_set_tile_shape")

    def set_overlap(self, on):
        check(lib().cbns_set_overlap(self._h, int(bool(on))), "cbns_set_overlap")

    def set_viscous_workgroups(self, n):
        """With overlap: workgroups of the side-stream viscous kernel (0 = one per unit)."""
        check(lib().cbns_set_viscous_workgroups(self._h, int(n)), "cbns_set_viscous_workgroups")

    def set_profiling(self, on):
        check(lib().cbns_set_profiling(self._h, int(bool(on))), "cbns_set_profiling")

    def set_profiling_classes(self, names):
        """Time only the launches of the named kernel classes (see kernel_times)."""
        mask = 0
        for n in names:
            mask |= 1 << self.KERNEL_CLASSES.index(n)
        check(lib().cbns_set_profiling_classes(self._h, mask), "cbns_set_profiling_classes")

    def kernel_times(self):
        """{class: (total_ms, launches)} of launches timed since profiling was turned on."""
        ms = (_D * 8)()
        n = (ctypes.c_int64 * 8)()
        check(lib().cbns_kernel_times(self._h, ms, n), "cbns_kernel_times")
        return {k: (ms[i], n[i]) for i, k in enumerate(self.KERNEL_CLASSES)}


def face_flux_eval(direction, gamma, cells, prim, precision=CBNS_FP64):
    """Device face flux on explicit inputs: cells (n, 6, 5), prim (n, 2, 5)."""
    cells = np.ascontiguousarray(cells, dtype=np.float64)
    prim = np.ascontiguousarray(prim, dtype=np.float64)
    n = cells.shape[0]
    out = np.empty((n, 5))
    check(lib().cbns_face_flux_eval(precision, direction, gamma, n, _p(cells), _p(prim), _p(out)),
          "cbns_face_flux_eval")
    return out


def face_flux_eval_noalloc(direction, gamma, cells, prim, precision=CBNS_FP64):
    """face_flux_eval staged through the library's fixed per-GPU scratch
    (cbns_face_flux_eval_noalloc): no device allocation, at most
    CBNS_NOALLOC_MAX_FACES faces per call."""
    cells = np.ascontiguousarray(cells, dtype=np.float64)
    prim = np.ascontiguousarray(prim, dtype=np.float64)
    n = cells.shape[0]
    out = np.empty((n, 5))
    check(lib().cbns_face_flux_eval_noalloc(precision, direction, gamma, n, _p(cells), _p(prim), _p(out)),
          "cbns_face_flux_eval_noalloc")
    return out


def batch_bytes_needed(params, dims=(1, 1, 1)):
    v = ctypes.c_uint64()
    check(lib().cbns_batch_bytes_needed(ctypes.byref(params), (_I * 3)(*dims), ctypes.byref(v)),
          "cbns_batch_bytes_needed")
    return v.value


class Batch:
    """Many independent solves advanced together (persistent kernels over a
    packed block table, one hipGraph per RK3 step)."""

    def __init__(self, budget_bytes, capacity=256):
        h = _VP()
        check(lib().cbns_batch_create(int(budget_bytes), int(capacity), ctypes.byref(h)), "cbns_batch_create")
        self._h = h
        self._N = {}

    def close(self):
        if getattr(self, "_h", None):
            lib().cbns_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add(self, params, dims=(1, 1, 1), precision=CBNS_FP64, Q=None, t_cur=0.0, mu0=0.0, steps=1 << 30):
        """Admit a solve; returns its id.  Raises CbnsError(-3) when the budget
        or capacity is full."""
        qp = None
        if Q is not None:
            Q = np.ascontiguousarray(Q, dtype=np.float64)
            assert Q.size == 5 * params.IM * params.JM * params.KM
            qp = _p(Q)
        i = _I()
        check(lib().cbns_batch_add(self._h, ctypes.byref(params), (_I * 3)(*dims), precision, qp, t_cur, mu0,
                                   int(steps), ctypes.byref(i)), "cbns_batch_add")
        self._N[i.value] = (params.IM, params.JM, params.KM)
        return i.value

    def remove(self, i):
        check(lib().cbns_batch_remove(self._h, i), "cbns_batch_remove")
        self._N.pop(i, None)

    def run(self, nsteps):
        check(lib().cbns_batch_run(self._h, int(nsteps)), "cbns_batch_run")

    def set_overlap(self, viscous_workgroups):
        """0: viscous kernels in line; n > 0: beside the face sweeps on a side
        stream with n persistent workgroups."""
        check(lib().cbns_batch_set_overlap(self._h, int(viscous_workgroups)), "cbns_batch_set_overlap")

    def set_band(self, rows):
        """Rows per band in the work order of the sweeps and the stage tail."""
        check(lib().cbns_batch_set_band(self._h, int(rows)), "cbns_batch_set_band")

    def set_workgroups(self, n):
        """Persistent workgroups of the memory-bound stage kernels (0, the
        default: a direct launch of one workgroup per tile for the stage tail
        and ghost fill, and the resident workgroups for the dt pass)."""
        check(lib().cbns_batch_set_workgroups(self._h, int(n)), "cbns_batch_set_workgroups")

    def sync(self):
        check(lib().cbns_batch_sync(self._h), "cbns_batch_sync")

    def busy(self):
        """Whether the last run's steps are still on the device (no wait)."""
        v = _I()
        check(lib().cbns_batch_query(self._h, ctypes.byref(v)), "cbns_batch_query")
        return bool(v.value)

    @property
    def pooled_bytes(self):
        """Device bytes kept from removed solves for reuse."""
        v = ctypes.c_uint64()
        check(lib().cbns_batch_pooled(self._h, ctypes.byref(v)), "cbns_batch_pooled")
        return v.value

    def trim(self):
        """Release the pooled arrays of removed solves."""
        check(lib().cbns_batch_trim(self._h), "cbns_batch_trim")

    def set_budgets(self, budgets):
        """{id: steps} granted until the next call."""
        ids = (_I * len(budgets))(*budgets.keys())
        st = (_I * len(budgets))(*budgets.values())
        check(lib().cbns_batch_set_budgets(self._h, ids, st, len(budgets)), "cbns_batch_set_budgets")

    def status(self, i):
        a, n, nf = _I(), _I(), _I()
        t = _D()
        check(lib().cbns_batch_status(self._h, i, ctypes.byref(a), ctypes.byref(n), ctypes.byref(t),
                                      ctypes.byref(nf)), "cbns_batch_status")
        return {"active": bool(a.value), "steps_done": n.value, "t_cur": t.value, "nonfinite": bool(nf.value)}

    def poll(self, max_ids=4096):
        ids = (_I * max_ids)()
        nd, nl = _I(), _I()
        check(lib().cbns_batch_poll(self._h, ids, max_ids, ctypes.byref(nd), ctypes.byref(nl)), "cbns_batch_poll")
        return [ids[k] for k in range(nd.value)], nl.value

    def poll_status(self, max_ids=4096):
        """Done solves with their status, in one read-back:
        ([{"id", "t_cur", "steps_done", "nonfinite"}], live)."""
        ids, ns, nf = (_I * max_ids)(), (_I * max_ids)(), (_I * max_ids)()
        t = (_D * max_ids)()
        nd, nl = _I(), _I()
        check(lib().cbns_batch_poll_status(self._h, ids, t, ns, nf, max_ids, ctypes.byref(nd), ctypes.byref(nl)),
              "cbns_batch_poll_status")
        return [{"id": ids[k], "t_cur": t[k], "steps_done": ns[k], "nonfinite": bool(nf[k])}
                for k in range(nd.value)], nl.value

    def get_state(self, i):
        N = self._N[i]
        out = np.empty((N[2], N[1], N[0], 5))
        check(lib().cbns_batch_get_state(self._h, i, _p(out)), "cbns_batch_get_state")
        return out

    @property
    def memory(self):
        u, b = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().cbns_batch_memory(self._h, ctypes.byref(u), ctypes.byref(b)), "cbns_batch_memory")
        return u.value, b.value
