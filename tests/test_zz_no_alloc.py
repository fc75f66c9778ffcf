"""lbm_get_fields_noalloc (fields staged through the library's fixed device array, no allocation)
against lbm_get_fields, through the C ABI: bit for bit on two calls in a row, the second with a
smaller input, in plain and reference modes; sizes above the array's maximum are rejected."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(lib, s):
    out = [np.full(s.shape, np.nan) for _ in range(4)]
    rc = lib.lbm_get_fields_noalloc(s.h, *(o.ctypes.data_as(C.c_void_p) for o in out))
    return rc, out


def _cfg(L, Q, n, mode, prec):
    nx, ny, nz = n
    kw = dict(tau0=0.06, u0=0.05, precision=L.PRECISIONS[prec], body=L.BODY_SPHERE,
              center=(nx / 3.0, ny / 2.0 - 0.5, nz / 2.0 - 0.5), r2=6.0, force_interval=2)
    if mode == "plain":
        return L.make_config(Q, nx, ny, nz, bc_mode=L.BC_PLAIN, perturb=1e-3, **kw)
    return L.make_config(Q, nx, ny, nz, bc_mode=L.BC_REFERENCE, **kw)


@pytest.mark.parametrize("first,second", [
    ((27, (64, 40, 36), "plain", "fp32"), (9, (50, 30, 1), "plain", "fp16s")),
    ((27, (40, 20, 16), "ref", "fp64"), (9, (30, 14, 1), "ref", "fp64")),
    ((27, (48, 24, 20), "plain", "fp64"), (27, (20, 12, 10), "ref", "fp64")),
])
def test_noalloc_fields_equal_the_allocating_call(pkg, first, second):
    L = pkg.lbm
    lib = L.lib()
    for Q, n, mode, prec in (first, second):        # the second input is the smaller one
        print(f"[noalloc] Q={Q} {n} {mode} {prec}", flush=True)
        with pkg.Solver(_cfg(L, Q, n, mode, prec)) as s:
            s.initialize()
            s.step(7)
            rc, out = _noalloc(lib, s)
            assert rc == 0, lib.lbm_last_error(s.h)
            ref = s.fields()
            for a, b in zip(out, ref):
                assert np.array_equal(a, b)
            s.step(3)                                # again on the same context, after more steps
            rc, out2 = _noalloc(lib, s)
            assert rc == 0
            for a, b in zip(out2, s.fields()):
                assert np.array_equal(a, b)
            assert not np.array_equal(out2[1], out[1])


def test_noalloc_rejects_sizes_above_its_maximum(pkg):
    L = pkg.lbm
    lib = L.lib()
    g = L.make_config(27, 512, 128, 65, tau0=0.06, u0=0.05, precision=L.FP16S, bc_mode=L.BC_PLAIN)  # 4 259 840 nodes
    with pkg.Solver(g) as s:
        s.initialize()
        s.step(1)
        rc, _ = _noalloc(lib, s)
        assert rc == 1
        assert b"maximum" in lib.lbm_last_error(s.h)
        p, u, v, w = s.fields()                     # the allocating call still serves it
        assert np.isfinite(u).all()
