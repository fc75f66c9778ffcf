"""The C-ABI library and the drop-in host driver, without a GPU.

  * liblbm_hip.so loads and exports every function include/lbm_capi.h declares;
  * configuration validation rejects bad input with LBM_ERR_INVALID before any HIP call;
  * lbm_config_default and the System parser reproduce the reference's derived parameters
    (3D MPI+KOKKOS+MPIIO/System.cpp:5-23, input.in; MPI Kokkos/System.cpp:3-18);
  * the velocity set matches the oracle's table index for index.
"""
import ctypes as C
import math
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, PKG_DIR

HEADER = os.path.join(ROOT, "include", "lbm_capi.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:extern\s+)?[\w\s\*]+?\b(lbm_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol(pkg):
    lib = pkg.lbm.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", pkg.lbm.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (lbm_\w+)", out))
    assert set(names) <= exported


def test_lattice_table_matches_oracle(pkg, orc):
    for Q in (9, 27):
        e, w, opp = pkg.lbm.lattice_table(Q)
        e2, w2, bb2 = orc.lattice(Q)
        assert np.array_equal(e, e2) and np.array_equal(w, w2) and np.array_equal(opp, bb2)
    with pytest.raises(ValueError):
        pkg.lbm.lattice_table(19)


def test_defaults_follow_reference_parameters(pkg):
    L = pkg.lbm
    c = L.default_config(27)
    # input.in: rho=1, R=8, Re=100, u0=0.1, 512x128x128; tau = u0 R / Re / cs2
    assert (c.nx, c.ny, c.nz) == (512, 128, 128)
    assert c.tau0 == 0.1 * 8.0 / 100.0 / (1.0 / 3.0)
    assert (c.cx, c.cy, c.cz) == (63.0, 63.0, 63.0) and c.r2 == (511 / 64.0) ** 2
    assert c.bc_mode == L.BC_REFERENCE and c.precision == L.FP64 and c.seed == 5374857
    c2 = L.default_config(9)
    assert (c2.nx, c2.ny) == (512, 128) and c2.tau0 == 0.1 * 12 / 100.0 / (1.0 / 3.0)
    assert c2.cx == 255.5 and c2.cy == 63.5 and c2.r2 == (0.03125 * 511 / 2.0) ** 2
    assert c2.force_interval == 1


@pytest.mark.parametrize("field,value", [("lattice", 19), ("nx", 2), ("tau0", 0.0), ("tau0", float("nan")),
                                         ("precision", 7), ("bc_mode", 5), ("ncoef", 9), ("force_interval", -1)])
def test_invalid_configs_are_rejected_before_any_gpu_call(pkg, field, value):
    L = pkg.lbm
    c = L.make_config(27, 16, 16, 16, tau0=0.1)
    setattr(c, field, value)
    h = C.c_void_p()
    rc = L.lib().lbm_create(C.byref(c), C.byref(h))
    assert rc == 1 and not h.value
    assert L.lib().lbm_last_error(None)


def test_reference_mode_requires_fp64(pkg):
    L = pkg.lbm
    c = L.make_config(27, 16, 16, 16, tau0=0.1, bc_mode=L.BC_REFERENCE, precision=L.FP32)
    h = C.c_void_p()
    assert L.lib().lbm_create(C.byref(c), C.byref(h)) == 1
    assert b"FP64" in L.lib().lbm_last_error(None)


def test_null_handles_are_errors_not_crashes(pkg):
    lib = pkg.lbm.lib()
    assert lib.lbm_step(None, 1) != 0
    assert lib.lbm_initialize(None) != 0
    assert lib.lbm_sync(None) != 0
    for stage in (lib.lbm_collide, lib.lbm_boundary, lib.lbm_stream, lib.lbm_update):
        assert stage(None) == 1          # LBM_ERR_INVALID
    lib.lbm_destroy(None)


def host_driver():
    exe = os.path.join(PKG_DIR, "host", "lbm_main")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", PKG_DIR, "host/lbm_main"], check=True)
    return exe


def test_system_parses_input_in_like_the_reference(tmp_path):
    exe = host_driver()
    (tmp_path / "input.in").write_text("1.0  8.0  100\n0.1  100  10\n512  128  128\n=============\nrho R Re\n")
    out = subprocess.run([exe, "--check"], cwd=tmp_path, capture_output=True, text=True, check=True).stdout
    vals = dict(re.findall(r"^(\w+)\s*=(\S+)$", out, flags=re.M))
    miu = 1.0 * 0.1 * 8.0 / 100
    assert float(vals["miu"]) == pytest.approx(miu)
    assert float(vals["tau"]) == pytest.approx(0.1 * 8.0 / 100 * 3)
    assert int(vals["Time"]) == int(math.pow(16.0, 2) / miu) and int(vals["inter"]) == int(vals["Time"]) // 10
    assert (vals["nx"], vals["ny"], vals["nz"]) == ("512", "128", "128")
    assert float(vals["Ma"]) == pytest.approx(0.1 / math.sqrt(1 / 3), rel=1e-5)


def test_system_2d_hard_coded_parameters():
    out = subprocess.run([host_driver(), "--2d", "--check"], capture_output=True, text=True, check=True).stdout
    vals = dict(re.findall(r"^(\w+)\s*=(\S+)$", out, flags=re.M))
    assert vals["Time"] == "1000" and vals["inter"] == "100"
    assert float(vals["tau"]) == pytest.approx(0.1 * 12 / 100 * 3)


def test_system_2d_single_node_parameters():
    """Kokkos/System.cpp:3-18: the MPI variant's values with Time 100000 and inter 10000."""
    out = subprocess.run([host_driver(), "--2d-single", "--check"], capture_output=True, text=True, check=True).stdout
    vals = dict(re.findall(r"^(\w+)\s*=(\S+)$", out, flags=re.M))
    assert vals["Time"] == "100000" and vals["inter"] == "10000"
    assert float(vals["tau"]) == pytest.approx(0.1 * 12 / 100 * 3)


def test_single_node_mode_is_d2q9_fp64_only(pkg):
    L = pkg.lbm
    for Q, prec in ((27, L.FP64), (9, L.FP32)):
        c = L.make_config(Q, 20, 10, 8 if Q == 27 else 1, tau0=0.036, u0=0.1, precision=prec,
                          bc_mode=L.BC_REFERENCE_SN, body=L.BODY_SPHERE, center=(4.0, 5.0, 0.0), r2=2.0)
        h = C.c_void_p()
        assert L.lib().lbm_create(C.byref(c), C.byref(h)) == 1 and not h.value   # LBM_ERR_INVALID
        msg = L.lib().lbm_last_error(None)
        assert (b"lattice 9 only" in msg) if Q == 27 else (b"FP64" in msg)


def test_legendre_basis_matches_oracle_inlet(pkg, orc):
    from cylinder_flow_amd.autograd import legendre_basis
    coef = pkg.configs.legendre_coefficients()
    assert abs(coef).sum() <= 0.05 + 1e-15 and coef[0] > 0
    ny, nz = 12, 10
    cfg = orc.make_cfg(27, 8, ny, nz, tau0=0.1, u0=0.05, inlet=orc.INLET_LEGENDRE, coef=tuple(coef))
    uin = orc.PlainState(cfg).inlet()
    phi = legendre_basis(ny, nz).numpy()
    assert np.allclose(np.einsum("k,kzy->zy", coef, phi), uin, atol=1e-15)


def test_strouhal_of_a_sine(pkg):
    t = np.arange(4000)
    cl = 0.3 * np.sin(2 * np.pi * t / 250.0 + 0.4)
    # period 250 steps sampled every step: St = D / (T u) with D = 20, u = 0.05
    assert abs(pkg.configs.strouhal(cl, 1.0, 20.0, 0.05) - 20.0 / (250.0 * 0.05)) < 1e-3
    # sampled every 10 steps, with 5 % noise: still the same frequency
    rng = np.random.default_rng(1)
    t10 = np.arange(0, 40000, 10)
    cl10 = np.sin(2 * np.pi * t10 / 2500.0) + 0.05 * rng.standard_normal(t10.size)
    assert abs(pkg.configs.strouhal(cl10, 10.0, 20.0, 0.05) / (20.0 / (2500.0 * 0.05)) - 1) < 0.02


def test_strouhal_band_removes_a_faster_mode(pkg):
    # shedding at St 0.22 under a stronger mode at St 0.9 (a channel's acoustic standing wave)
    t = np.arange(0, 60000, 10)
    D, u = 40.0, 0.05
    cl = 0.2 * np.sin(2 * np.pi * t * 0.22 * u / D) + 0.3 * np.sin(2 * np.pi * t * 0.9 * u / D)
    assert abs(pkg.configs.strouhal(cl, 10.0, D, u, band=(0.05, 0.5)) / 0.22 - 1) < 0.02


def test_strouhal_is_nan_without_shedding(pkg):
    rng = np.random.default_rng(2)
    assert np.isnan(pkg.configs.strouhal(1e-3 * rng.standard_normal(2000), 10.0, 40.0, 0.05))
    assert np.isnan(pkg.configs.strouhal(np.sin(np.arange(100) / 30.0), 10.0, 40.0, 0.05))   # < 3 periods
