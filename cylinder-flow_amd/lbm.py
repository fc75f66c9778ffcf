"""ctypes binding of liblbm_hip.so (include/lbm_capi.h) for tests, benchmarks and scripts.

The compute path is the HIP library; this module only marshals arguments.  Loading fails
loudly when the library has not been built -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbm_hip.so")

FP64, FP32, FP16S, BF16S = 0, 1, 2, 3
BC_REFERENCE, BC_PLAIN, BC_REFERENCE_SN = 0, 1, 2
BODY_NONE, BODY_SPHERE, BODY_ZCYL, BODY_MASK = 0, 1, 2, 3
INLET_UNIFORM, INLET_PARABOLIC, INLET_LEGENDRE, INLET_PROFILE = 0, 1, 2, 3
PRECISIONS = {"fp64": FP64, "fp32": FP32, "fp16s": FP16S, "bf16s": BF16S}
STORAGE_BYTES = {FP64: 8, FP32: 4, FP16S: 2, BF16S: 2}


class LbmConfig(C.Structure):
    _fields_ = [
        ("lattice", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("tau0", C.c_double), ("rho0", C.c_double), ("u0", C.c_double),
        ("precision", C.c_int), ("bc_mode", C.c_int), ("periodic", C.c_int * 3),
        ("body", C.c_int), ("cx", C.c_double), ("cy", C.c_double), ("cz", C.c_double),
        ("r2", C.c_double),
        ("inlet", C.c_int), ("ncoef", C.c_int), ("coef", C.c_double * 8),
        ("seed", C.c_uint64), ("perturb", C.c_double),
        ("force_interval", C.c_int), ("force_limit", C.c_double), ("store_uref", C.c_double),
    ]


class LbmInfo(C.Structure):
    _fields_ = [
        ("lattice", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("pitch", C.c_int),
        ("nodes", C.c_size_t), ("state_bytes", C.c_size_t), ("bytes_per_update", C.c_double),
        ("step", C.c_long), ("solid_nodes", C.c_int), ("boundary_nodes", C.c_int),
        ("nodes_per_lane", C.c_int),
    ]


class LbmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"liblbm_hip error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() (make -C cylinder-flow_amd)")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        I = C.c_int
        sig = {
            "lbm_config_default": (I, [C.POINTER(LbmConfig), I]),
            "lbm_lattice_table": (I, [I, P, P, P]),
            "lbm_create": (I, [C.POINTER(LbmConfig), C.POINTER(P)]),
            "lbm_destroy": (None, [P]),
            "lbm_last_error": (C.c_char_p, [P]),
            "lbm_set_mask": (I, [P, P]),
            "lbm_set_solid_fraction": (I, [P, P]),
            "lbm_set_inlet_profile": (I, [P, P]),
            "lbm_set_perturbation": (I, [P, P]),
            "lbm_set_ran": (I, [P, P, C.c_size_t]),
            "lbm_initialize": (I, [P]),
            "lbm_step": (I, [P, I]),
            "lbm_step_timed": (I, [P, I, C.POINTER(C.c_double)]),
            "lbm_sync": (I, [P]),
            "lbm_get_fields_noalloc": (I, [P, P, P, P, P]),
            "lbm_set_step": (I, [P, C.c_long]),
            "lbm_collide": (I, [P]),
            "lbm_boundary": (I, [P]),
            "lbm_stream": (I, [P]),
            "lbm_update": (I, [P]),
            "lbm_get_fields": (I, [P, P, P, P, P]),
            "lbm_get_populations": (I, [P, P]),
            "lbm_set_populations": (I, [P, P]),
            "lbm_get_forces": (I, [P, P, I, C.POINTER(I)]),
            "lbm_info_get": (I, [P, C.POINTER(LbmInfo)]),
            "lbm_snapshot_async": (I, [P, C.c_char_p, I]),
            "lbm_step_kernel_timed": (I, [P, I, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "lbm_snapshot_wait": (I, [P]),
            "lbm_adjoint": (I, [P, I, I, C.c_double, I, P, P, P, P, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def lattice_table(Q):
    e = np.zeros((Q, 3), np.int32)
    w = np.zeros(Q)
    opp = np.zeros(Q, np.int32)
    if lib().lbm_lattice_table(Q, _p(e), _p(w), _p(opp)) != 0:
        raise ValueError("lattice must be 9 or 27")
    return e, w, opp


def default_config(Q) -> LbmConfig:
    c = LbmConfig()
    if lib().lbm_config_default(C.byref(c), Q) != 0:
        raise ValueError("lattice must be 9 or 27")
    return c


def make_config(Q, nx, ny, nz=1, *, tau0, u0=0.05, precision=FP64, bc_mode=BC_PLAIN,
                periodic=(False, False, False), body=BODY_NONE, center=(0.0, 0.0, 0.0), r2=0.0,
                inlet=INLET_UNIFORM, coef=(), seed=5374857, perturb=0.0, force_interval=0,
                force_limit=0.0, store_uref=0.0) -> LbmConfig:
    c = LbmConfig()
    c.lattice, c.nx, c.ny, c.nz = Q, nx, ny, (nz if Q == 27 else 1)
    c.tau0, c.rho0, c.u0 = tau0, 1.0, u0
    c.precision, c.bc_mode = precision, bc_mode
    for k in range(3):
        c.periodic[k] = int(bool(periodic[k]))
    c.body = body
    c.cx, c.cy, c.cz = center
    c.r2 = r2
    c.inlet, c.ncoef = inlet, len(coef)
    for k, v in enumerate(coef):
        c.coef[k] = v
    c.seed, c.perturb = seed, perturb
    c.force_interval, c.force_limit, c.store_uref = force_interval, force_limit, store_uref
    return c


class Solver:
    """One lbm_ctx: the drop-in for the reference's System + LBM pair."""

    def __init__(self, cfg: LbmConfig):
        self.cfg = cfg
        h = C.c_void_p()
        rc = lib().lbm_create(C.byref(cfg), C.byref(h))
        if rc != 0:
            raise LbmError(rc, lib().lbm_last_error(None).decode())
        self.h = h
        self.Q = cfg.lattice
        self.shape = (cfg.nz if cfg.lattice == 27 else 1, cfg.ny, cfg.nx)
        self.N = int(np.prod(self.shape))

    def _chk(self, rc):
        if rc != 0:
            raise LbmError(rc, lib().lbm_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            lib().lbm_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # set-up
    def set_mask(self, solid):
        a = np.ascontiguousarray(solid, np.uint8).reshape(-1)
        assert a.size == self.N
        self._chk(lib().lbm_set_mask(self.h, _p(a)))

    def set_solid_fraction(self, s):
        a = np.ascontiguousarray(s, np.float64).reshape(-1)
        assert a.size == self.N
        self._chk(lib().lbm_set_solid_fraction(self.h, _p(a)))

    def set_inlet_profile(self, uin):
        a = np.ascontiguousarray(uin, np.float64).reshape(-1)
        assert a.size == self.shape[0] * self.shape[1]
        self._chk(lib().lbm_set_inlet_profile(self.h, _p(a)))

    def set_perturbation(self, xi):
        a = np.ascontiguousarray(xi, np.float64).reshape(-1)
        assert a.size == self.N
        self._chk(lib().lbm_set_perturbation(self.h, _p(a)))

    def set_ran(self, ran):
        a = np.ascontiguousarray(ran, np.int32).reshape(-1)
        self._chk(lib().lbm_set_ran(self.h, _p(a), a.size))

    def initialize(self):
        self._chk(lib().lbm_initialize(self.h))
        return self

    # time stepping
    def step(self, n=1):
        self._chk(lib().lbm_step(self.h, n))

    def step_timed(self, n):
        ms = C.c_double()
        self._chk(lib().lbm_step_timed(self.h, n, C.byref(ms)))
        return ms.value

    def step_kernel_timed(self, n):
        """(ms of the n steps, ms summed over the steps of the bulk kernel alone)"""
        t, k = C.c_double(), C.c_double()
        self._chk(lib().lbm_step_kernel_timed(self.h, n, C.byref(t), C.byref(k)))
        return t.value, k.value

    def sync(self):
        self._chk(lib().lbm_sync(self.h))

    # one step in the reference's stages (reference boundary modes), in this order
    def collide(self):
        self._chk(lib().lbm_collide(self.h))

    def boundary(self):
        self._chk(lib().lbm_boundary(self.h))

    def stream(self):
        self._chk(lib().lbm_stream(self.h))

    def update(self):
        self._chk(lib().lbm_update(self.h))

    # data out
    def fields(self):
        out = [np.zeros(self.shape) for _ in range(4)]
        self._chk(lib().lbm_get_fields(self.h, *(_p(a) for a in out)))
        return tuple(out)  # p, u, v, w

    def populations(self):
        out = np.zeros((self.Q,) + self.shape)
        self._chk(lib().lbm_get_populations(self.h, _p(out)))
        return out

    def set_populations(self, f):
        a = np.ascontiguousarray(f, np.float64)
        assert a.size == self.Q * self.N
        self._chk(lib().lbm_set_populations(self.h, _p(a)))

    def set_step(self, step):
        self._chk(lib().lbm_set_step(self.h, int(step)))

    def forces(self):
        n = C.c_int()
        self._chk(lib().lbm_get_forces(self.h, None, 0, C.byref(n)))
        out = np.zeros((n.value, 4))
        if n.value:
            self._chk(lib().lbm_get_forces(self.h, _p(out), n.value, C.byref(n)))
        return out

    def info(self) -> LbmInfo:
        i = LbmInfo()
        self._chk(lib().lbm_info_get(self.h, C.byref(i)))
        return i

    def snapshot(self, path, n):
        self._chk(lib().lbm_snapshot_async(self.h, str(path).encode(), n))

    def snapshot_wait(self):
        self._chk(lib().lbm_snapshot_wait(self.h))

    def adjoint(self, n0, n1, scale=1.0, checkpoint_every=0):
        """Returns (J, dJ/ds (shape), dJ/du_in (nz, ny), ms_forward, ms_adjoint)."""
        J = np.zeros(1)
        ds = np.zeros(self.shape)
        du = np.zeros(self.shape[:2])
        tf = np.zeros(1)
        ta = np.zeros(1)
        self._chk(lib().lbm_adjoint(self.h, n0, n1, scale, checkpoint_every, _p(J), _p(ds), _p(du),
                                    _p(tf), _p(ta)))
        return J[0], ds, du, tf[0], ta[0]
