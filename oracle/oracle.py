"""ctypes binding of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py``, which compare the product against it.  The product
library never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# LBM_ORACLE_LIB selects another build of the same sources (the sanitiser build, or a
# deliberately broken copy that the pin tests must reject)
LIB_PATH = os.environ.get("LBM_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

BODY_NONE, BODY_SPHERE, BODY_ZCYL, BODY_MASK = 0, 1, 2, 3
INLET_UNIFORM, INLET_PARABOLIC, INLET_LEGENDRE, INLET_PROFILE = 0, 1, 2, 3


class OrcCfg(C.Structure):
    _fields_ = [
        ("Q", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("tau0", C.c_double), ("u0", C.c_double),
        ("body", C.c_int), ("cx", C.c_double), ("cy", C.c_double), ("cz", C.c_double),
        ("r2", C.c_double),
        ("inlet", C.c_int), ("ncoef", C.c_int), ("coef", C.c_double * 8),
        ("seed", C.c_uint64), ("perturb", C.c_double),
        ("xper", C.c_int), ("yper", C.c_int), ("zper", C.c_int),
        ("force_interval", C.c_int), ("variant", C.c_int),
    ]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-B", "-C", HERE, "liboracle.so"], check=True)


def f32_lib():
    """The float build (liboracle_f32.so: populations, fields and step arithmetic of both modes
    in float, no adjoint), for the CPU baseline's FP32 legs.  Same create/step interface."""
    path = os.path.join(HERE, "liboracle_f32.so")
    if not os.path.exists(path):
        subprocess.run(["make", "-s", "-B", "-C", HERE, "liboracle_f32.so"], check=True)
    L = C.CDLL(path)
    P = C.c_void_p
    L.orc_plain_create.restype = P
    L.orc_plain_create.argtypes = [C.POINTER(OrcCfg), P, P, P, P]
    L.orc_plain_destroy.argtypes = [P]
    L.orc_plain_step.argtypes = [P, C.c_int]
    L.orc_plain_fields.argtypes = [P, P, P, P, P]
    L.orc_plain_populations.argtypes = [P, P]
    L.orc_plain_forces.restype = C.c_int
    L.orc_plain_forces.argtypes = [P, P, C.c_int]
    L.orc_ref_create.restype = P
    L.orc_ref_create.argtypes = [C.POINTER(OrcCfg), P]
    L.orc_ref_destroy.argtypes = [P]
    L.orc_ref_step.argtypes = [P, C.c_int]
    L.orc_ref_populations.argtypes = [P, P]
    L.orc_num_threads.restype = C.c_int
    L.orc_set_threads.argtypes = [C.c_int]
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        L.orc_hash.restype = C.c_uint64
        L.orc_hash.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_xi.restype = C.c_double
        L.orc_xi.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_ran100.restype = C.c_int
        L.orc_ran100.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_lattice.argtypes = [C.c_int, P, P, P]
        L.orc_feq.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, P]
        L.orc_ref_create.restype = P
        L.orc_ref_create.argtypes = [C.POINTER(OrcCfg), P]
        L.orc_ref_destroy.argtypes = [P]
        L.orc_ref_extents.argtypes = [P] + [C.POINTER(C.c_int)] * 4
        L.orc_ref_step.argtypes = [P, C.c_int]
        L.orc_ref_stage.argtypes = [P, C.c_int]
        L.orc_ref_fields.argtypes = [P, P, P, P, P]
        L.orc_ref_populations.argtypes = [P, P]
        L.orc_ref_raw.restype = C.POINTER(C.c_double)
        L.orc_ref_raw.argtypes = [P]
        L.orc_ref_set_raw.argtypes = [P, P]
        for name in ("collision", "boundary_fills", "usrbc", "stream", "update"):
            getattr(L, "orc_ref_" + name).argtypes = [P]
        L.orc_ref_mem.argtypes = [P, P]
        L.orc_ref_forces.restype = C.c_int
        L.orc_ref_forces.argtypes = [P, P, C.c_int]
        L.orc_plain_create.restype = P
        L.orc_plain_create.argtypes = [C.POINTER(OrcCfg), P, P, P, P]
        L.orc_plain_destroy.argtypes = [P]
        L.orc_plain_step.argtypes = [P, C.c_int]
        L.orc_plain_fields.argtypes = [P, P, P, P, P]
        L.orc_plain_populations.argtypes = [P, P]
        L.orc_plain_set_populations.argtypes = [P, P]
        L.orc_plain_last_g.argtypes = [P, P]
        L.orc_plain_inlet.argtypes = [P, P]
        L.orc_plain_forces.restype = C.c_int
        L.orc_plain_forces.argtypes = [P, P, C.c_int]
        L.orc_num_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_adjoint.restype = C.c_double
        L.orc_adjoint.argtypes = [C.POINTER(OrcCfg), P, P, P, P, C.c_int, C.c_int, C.c_int,
                                  C.c_double, C.c_int, P, P]
        L.orc_objective.restype = C.c_double
        L.orc_objective.argtypes = [C.POINTER(OrcCfg), P, P, P, P, C.c_int, C.c_int, C.c_int,
                                    C.c_double]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def make_cfg(Q, nx, ny, nz=1, tau0=0.03, u0=0.05, body=BODY_NONE, center=(0.0, 0.0, 0.0),
             r2=0.0, inlet=INLET_UNIFORM, coef=(), seed=5374857, perturb=0.0,
             periodic=(False, False, False), force_interval=0, variant=0) -> OrcCfg:
    c = OrcCfg()
    c.Q, c.nx, c.ny, c.nz = Q, nx, ny, (nz if Q == 27 else 1)
    c.tau0, c.u0 = tau0, u0
    c.body = body
    c.cx, c.cy, c.cz = center
    c.r2 = r2
    c.inlet = inlet
    c.ncoef = len(coef)
    for i, v in enumerate(coef):
        c.coef[i] = v
    c.seed, c.perturb = seed, perturb
    c.xper, c.yper, c.zper = (int(bool(p)) for p in periodic)
    c.force_interval = force_interval
    c.variant = variant
    return c


def lattice(Q):
    e = np.zeros((Q, 3), np.int32)
    w = np.zeros(Q, np.float64)
    bb = np.zeros(Q, np.int32)
    assert lib().orc_lattice(Q, _ptr(e), _ptr(w), _ptr(bb)) == 0
    return e, w, bb


def feq(Q, p, u, v=0.0, w=0.0):
    out = np.zeros(Q)
    assert lib().orc_feq(Q, p, u, v, w, _ptr(out)) == 0
    return out


class RefState:
    """Reference-mode oracle: the reference kernels in the reference order."""

    def __init__(self, cfg: OrcCfg, ran: np.ndarray | None = None):
        self.cfg = cfg
        self._ran = None if ran is None else np.ascontiguousarray(ran, np.int32)
        self.h = lib().orc_ref_create(C.byref(cfg), _ptr(self._ran))
        if not self.h:
            raise ValueError("orc_ref_create failed")
        lx, ly, lz, g = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        lib().orc_ref_extents(self.h, C.byref(lx), C.byref(ly), C.byref(lz), C.byref(g))
        self.lx, self.ly, self.lz, self.ghost = lx.value, ly.value, lz.value, g.value
        self.Q = cfg.Q
        self.shape = (cfg.nz, cfg.ny, cfg.nx)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_ref_destroy(self.h)
            self.h = None

    def step(self, n=1):
        lib().orc_ref_step(self.h, n)

    def stage(self, k):
        """0 Collision, 1 fills + usrbc (+ momentum exchange), 2 stream1 + stream2, 3 Update."""
        lib().orc_ref_stage(self.h, k)

    def fields(self):
        out = [np.zeros(self.shape) for _ in range(4)]
        lib().orc_ref_fields(self.h, *(_ptr(a) for a in out))
        return tuple(out)  # p, u, v, w

    def populations(self):
        out = np.zeros((self.Q,) + self.shape)
        lib().orc_ref_populations(self.h, _ptr(out))
        return out

    def raw(self):
        n = self.lx * self.ly * self.lz * self.Q
        return np.ctypeslib.as_array(lib().orc_ref_raw(self.h), (n,)).copy()

    def set_raw(self, f):
        f = np.ascontiguousarray(f, np.float64)
        lib().orc_ref_set_raw(self.h, _ptr(f))

    def forces(self):
        n = lib().orc_ref_forces(self.h, None, 0)
        out = np.zeros((n, 4))
        lib().orc_ref_forces(self.h, _ptr(out), n)
        return out


class PlainState:
    """Plain-mode oracle (push form)."""

    def __init__(self, cfg: OrcCfg, mask=None, sfrac=None, xi=None, uin=None):
        self.cfg = cfg
        self.Q = cfg.Q
        self.shape = (cfg.nz, cfg.ny, cfg.nx)
        self._keep = [None if a is None else np.ascontiguousarray(a, dt)
                      for a, dt in ((mask, np.uint8), (sfrac, np.float64), (xi, np.float64),
                                    (uin, np.float64))]
        self.h = lib().orc_plain_create(C.byref(cfg), *(_ptr(a) for a in self._keep))
        if not self.h:
            raise ValueError("orc_plain_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_plain_destroy(self.h)
            self.h = None

    def step(self, n=1):
        lib().orc_plain_step(self.h, n)

    def fields(self):
        out = [np.zeros(self.shape) for _ in range(4)]
        lib().orc_plain_fields(self.h, *(_ptr(a) for a in out))
        return tuple(out)

    def populations(self):
        out = np.zeros((self.Q,) + self.shape)
        lib().orc_plain_populations(self.h, _ptr(out))
        return out

    def set_populations(self, f):
        f = np.ascontiguousarray(f, np.float64)
        lib().orc_plain_set_populations(self.h, _ptr(f))

    def last_g(self):
        out = np.zeros((self.Q,) + self.shape)
        lib().orc_plain_last_g(self.h, _ptr(out))
        return out

    def inlet(self):
        out = np.zeros(self.shape[:2])
        lib().orc_plain_inlet(self.h, _ptr(out))
        return out

    def forces(self):
        n = lib().orc_plain_forces(self.h, None, 0)
        out = np.zeros((n, 4))
        lib().orc_plain_forces(self.h, _ptr(out), n)
        return out


def adjoint(cfg, mask=None, sfrac=None, xi=None, uin=None, n0=1, n1=10, scale=1.0, K=0):
    """FP64 CPU adjoint: returns (J, dJ/ds (nz,ny,nx), dJ/du_in (nz,ny))."""
    shape = (cfg.nz, cfg.ny, cfg.nx)
    keep = [None if a is None else np.ascontiguousarray(a, dt)
            for a, dt in ((mask, np.uint8), (sfrac, np.float64), (xi, np.float64),
                          (uin, np.float64))]
    ds = np.zeros(shape)
    du = np.zeros(shape[:2])
    J = lib().orc_adjoint(C.byref(cfg), *(_ptr(a) for a in keep), n1, n0, n1, scale, K,
                          _ptr(ds), _ptr(du))
    return J, ds, du


def objective(cfg, mask=None, sfrac=None, xi=None, uin=None, n0=1, n1=10, scale=1.0):
    keep = [None if a is None else np.ascontiguousarray(a, dt)
            for a, dt in ((mask, np.uint8), (sfrac, np.float64), (xi, np.float64),
                          (uin, np.float64))]
    return lib().orc_objective(C.byref(cfg), *(_ptr(a) for a in keep), n1, n0, n1, scale)
