"""The drop-in host layer on the GPU: the asynchronous Tecplot snapshot writer (lbm_snapshot_async /
lbm_snapshot_wait, the reference's LBM::MPIoutput) and the driver host/lbm_main running the
reference's main loop from an input.in (3D) or its hard-coded 2D case, checked against the same
configuration stepped through the Python binding of the same library.

Snapshot layout (3D MPI+KOKKOS+MPIIO/lbm.cpp:528-739, MPI Kokkos/lbm.cpp:548-737 with one rank):
a 332-byte (3D) / 276-byte (2D) header ending in the (min, max) pairs of u, v, (w,) p, then the
coordinates x = 4 i/(glx-1), y = j/(gly-1), (z = k/(glz-1)) and u, v, (w,) p as doubles in
x-fastest order."""
import os
import subprocess

import numpy as np
import pytest

from test_capi_host import host_driver

pytestmark = pytest.mark.gpu

HEADER = {3: 332, 2: 276}
DIMS_AT = {3: 156, 2: 140}        # byte offset of the int32 triple (glx, gly, glz|1)


def read_plt(path, D, shape):
    nz, ny, nx = shape
    N = nx * ny * nz
    raw = open(path, "rb").read()
    nvar = 7 if D == 3 else 5
    assert len(raw) == HEADER[D] + 8 * nvar * N
    assert raw[:8] == b"#!TDV112"
    assert tuple(np.frombuffer(raw, np.int32, 3, DIMS_AT[D])) == (nx, ny, nz if D == 3 else 1)
    nf = nvar - D
    mm = np.frombuffer(raw, np.float64, 2 * nf, HEADER[D] - 16 * nf).reshape(nf, 2)
    data = np.frombuffer(raw, np.float64, nvar * N, HEADER[D]).reshape(nvar, N)
    return mm, data


def check_snapshot(path, D, fields):
    """fields = (p, u, v, w) arrays of shape (nz, ny, nx) from lbm_get_fields."""
    p, u, v, w = fields
    nz, ny, nx = p.shape
    mm, data = read_plt(path, D, p.shape)
    kk, jj, ii = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    coords = [4.0 * ii / (nx - 1), jj / (ny - 1)] + ([kk / (nz - 1)] if D == 3 else [])
    for c, ref in zip(data[:D], coords):
        assert np.array_equal(c, ref.ravel())
    vars_ = (u, v, w, p) if D == 3 else (u, v, p)
    for k, f in enumerate(vars_):
        assert np.array_equal(data[D + k], f.ravel())
        assert (mm[k, 0], mm[k, 1]) == (f.min(), f.max())


@pytest.mark.parametrize("Q,prec,bc", [(27, "fp32", "plain"), (27, "fp64", "ref"), (9, "fp16s", "plain"),
                                       (9, "fp64", "ref")])
def test_snapshot_matches_fields(pkg, tmp_path, Q, prec, bc):
    L = pkg.lbm
    nx, ny, nz = (40, 24, 20) if Q == 27 else (96, 40, 1)
    c = L.make_config(Q, nx, ny, nz, tau0=0.05, u0=0.05, precision=L.PRECISIONS[prec],
                      bc_mode=L.BC_PLAIN if bc == "plain" else L.BC_REFERENCE, body=L.BODY_SPHERE,
                      center=(nx / 4, ny / 2 - 0.5, nz / 2 - 0.5), r2=9.0, perturb=1e-3)
    print(f"[host] snapshot D2Q{Q} {prec} {bc}", flush=True)
    with L.Solver(c) as s:
        s.initialize()
        s.step(7)
        s.snapshot(tmp_path / "a.plt", 7)
        s.step(5)                                   # keeps stepping while the file is written
        s.snapshot(tmp_path / "b.plt", 12)          # waits for the first writer
        s.snapshot_wait()
        f12 = s.fields()
    D = 3 if Q == 27 else 2
    check_snapshot(tmp_path / "b.plt", D, f12)
    mm, _ = read_plt(tmp_path / "a.plt", D, f12[0].shape)
    assert np.isfinite(mm).all()


def run_driver(args, cwd):
    print(f"[host] lbm_main {' '.join(args)}", flush=True)
    r = subprocess.run([host_driver()] + args, cwd=cwd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_driver_3d_input_in(pkg, tmp_path):
    """rho R Re / u0 Time inter / nx ny nz = 1 2 10 / 0.1 . . / 64 32 32: tau 0.06,
    Time = (2R)^2/miu = 800, inter = 80; --time 160 writes output0..2.plt."""
    (tmp_path / "input.in").write_text("1.0 2.0 10\n0.1 100 10\n64 32 32\n=============\nrho R Re\n")
    out = run_driver(["--time", "160"], tmp_path)
    assert out.count("time=") == 2
    L = pkg.lbm
    c = L.default_config(27)
    c.nx, c.ny, c.nz = 64, 32, 32
    c.tau0, c.u0 = 0.1 * 2.0 / 10 / (1.0 / 3.0), 0.1          # System.cpp: u0 R / Re / cs2
    c.cx = c.cy = c.cz = float(63 // 8)
    c.r2 = (63 / 64.0) ** 2
    with L.Solver(c) as s:
        s.initialize()
        f0 = s.fields()
        s.step(80)
        f1 = s.fields()
        s.step(80)
        f2 = s.fields()
    for n, f in enumerate((f0, f1, f2)):
        check_snapshot(tmp_path / f"output{n}.plt", 3, f)
    assert not (tmp_path / "output3.plt").exists()


def test_driver_2d_test_dat(pkg, tmp_path):
    """The 2D driver (512 x 128, Re 100, u0 0.1): Output every 100 steps as ASCII x y u v p and
    Cd, Cl = 2 F/(u0^2 D) with D = 0.03125 (nx - 1) appended to test.dat every step."""
    run_driver(["--2d", "--time", "200"], tmp_path)
    L = pkg.lbm
    with L.Solver(L.default_config(9)) as s:
        s.initialize()
        s.step(200)
        F = s.forces()
        p, u, v, _ = s.fields()
    cd = np.loadtxt(tmp_path / "test.dat")
    assert cd.shape == (len(F), 2) and len(F) >= 200
    scale = 2.0 / 0.1 ** 2 / (0.03125 * 511)
    assert np.allclose(cd, F[:, 1:3] * scale, rtol=1e-14, atol=1e-16)
    with open(tmp_path / "output20.dat") as fh:
        assert fh.readline().strip() == "variables=x,y,u,v,p"
        assert fh.readline().strip() == "zone I=512,J=128"
    a = np.loadtxt(tmp_path / "output20.dat", skiprows=2)
    assert a.shape == (512 * 128, 5)
    for k, f in enumerate((u, v, p)):
        assert np.abs(a[:, 2 + k] - f.ravel()).max() <= 1e-8
    assert os.path.exists(tmp_path / "output00.dat") and os.path.exists(tmp_path / "output10.dat")


def test_driver_2d_single_node_variant(pkg, tmp_path):
    """--2d-single: the single-node 2D variant's main (512 x 128 -> 513 x 129 nodes); Output(0)
    writes output0.dat with x = 4 i / (nx - 1), equal to the initial fields of the same
    configuration through the Python binding (LBM_BC_REFERENCE_SN)."""
    run_driver(["--2d-single", "--time", "30"], tmp_path)
    L = pkg.lbm
    c = L.make_config(9, 513, 129, tau0=0.1 * 12 / 100 / (1.0 / 3.0), u0=0.1, precision=L.FP64,
                      bc_mode=L.BC_REFERENCE_SN, body=L.BODY_SPHERE, center=(0.25 * 512 - 2.0, 0.5 * 128, 0.0),
                      r2=0.01 * 128 * 128, force_interval=0)
    with L.Solver(c) as s:
        s.initialize()
        p, u, v, _ = s.fields()
    with open(tmp_path / "output0.dat") as fh:
        assert fh.readline().strip() == "variables=x,y,u,v,p"
        assert fh.readline().strip() == "zone I=513,J=129"
    a = np.loadtxt(tmp_path / "output0.dat", skiprows=2)
    assert a.shape == (513 * 129, 5)
    assert np.allclose(a[:513, 0], 4.0 * np.arange(513) / 512, atol=5e-9)
    for k, f in enumerate((u, v, p)):
        assert np.abs(a[:, 2 + k] - f.ravel()).max() <= 1e-8
    assert not (tmp_path / "test.dat").exists() and not (tmp_path / "output00.dat").exists()


@pytest.mark.parametrize("mode", ["3d", "2d"])
def test_driver_staged_equals_fused(tmp_path, mode):
    """LBMOptions::staged (`staged = yes` in input.in, or --staged): Collision(), Streaming() and
    Update() each run their own stage through lbm_collide / lbm_boundary + lbm_stream /
    lbm_update; the snapshots and the test.dat drag record equal the fused run's byte for byte."""
    outs = []
    for staged in (False, True):
        d = tmp_path / ("staged" if staged else "fused")
        d.mkdir()
        if mode == "3d":
            (d / "input.in").write_text("1.0 2.0 10\n0.1 100 10\n48 24 24\n=============\nrho R Re\n"
                                        + ("staged = yes\n" if staged else ""))
            run_driver(["--time", "120"], d)
            names = ["output0.plt", "output1.plt"]
        else:
            run_driver(["--2d", "--time", "120"] + (["--staged"] if staged else []), d)
            names = ["output00.dat", "output10.dat", "test.dat"]
        outs.append([(d / n).read_bytes() for n in names])
    for a, b in zip(*outs):
        assert len(a) > 0 and a == b
