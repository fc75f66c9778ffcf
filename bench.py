#!/usr/bin/env python3
"""Benchmark: D3Q27 MLUPS and HBM-roofline fraction on one MI355X, FP16-shifted vs FP32 storage,
plus the adjoint/forward time ratio.  Prints ONE JSON line.

Workload (the largest single-GPU configuration): D3Q27 BGK, 512x256x256 nodes (33.5 M),
z-spanning cylinder D=40 at (128,128), Re=300, u=0.05, half-way bounce-back on the body and the
four side walls, Zou-He inlet (regularised), zero-gradient outlet, drag/lift every 10 steps.  Synthetic
initial state (seeded perturbation); the populations are resident in HBM before timing starts.

A "step" is one fused lattice update of the whole domain: the bounce-back copies into the solid
slots (k_link_pairs, or k_link_v with the momentum exchange on force steps), k_fast_v (packed
pull-collide-store over every node) and, when its list is not empty, k_fix_v, back to back on
the solver's stream.  MLUPS = nodes * steps / seconds over
exactly --steps steps bracketed by stream synchronisation.  The roofline is that of the
dominant kernel, k_fast_v: the algorithmic 2*Q*sizeof(storage) bytes per node update times the
nodes one launch updates, over its average launch duration measured with HIP events recorded
just before and after each launch on the solver's own stream (a second context with forces
off).  roofline.step gives the same for the whole step (events around the steps), and
roofline.traffic the counter-measured HBM bytes per launch (profiles/traffic.json).
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "cylinder-flow_amd")
HBM_PEAK_GBS = 8000.0
METRIC = "D3Q27 MLUPS + % HBM roofline (FP32 vs FP16-store); adjoint/forward time"


def load_package():
    if "cylinder_flow_amd" in sys.modules:
        return sys.modules["cylinder_flow_amd"]
    spec = importlib.util.spec_from_file_location("cylinder_flow_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["cylinder_flow_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


_hip = None


def device_sync():
    """hipDeviceSynchronize: drains every stream, the solver's included (the timing bracket; the
    solver's own lbm_sync also checks the state for NaN/Inf, which is not part of a step)."""
    global _hip
    if _hip is None:
        import ctypes as C
        try:
            _hip = C.CDLL("libamdhip64.so")
        except OSError:
            _hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    if _hip.hipDeviceSynchronize() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")


def run_leg(pkg, precision, steps, warmup, force_interval):
    W = pkg.configs.zcylinder3d_512x256x256(precision=precision, force_interval=force_interval)
    s = pkg.Solver(W.cfg).initialize()
    s.step(warmup)
    s.sync()
    device_sync()
    t0 = time.perf_counter()
    s.step(steps)
    device_sync()
    wall = time.perf_counter() - t0
    s.sync()                      # forces and state checked for NaN/Inf, after the timed steps
    info = s.info()
    return s, W, wall, info


def kernel_roofline(pkg, precision, steps, bytes_per_update, nodes):
    """Forces off, HIP events on the solver's stream: the average duration of the dominant
    kernel (k_fast_v, events recorded just before and after each launch) and of the whole step
    (k_link_pairs + k_fast_v, events around the steps, in a run without the per-kernel events).
    Returns (kernel seconds, kernel GB/s, step seconds, step GB/s) per launch / step."""
    W = pkg.configs.zcylinder3d_512x256x256(precision=precision, force_interval=0)
    s = pkg.Solver(W.cfg).initialize()
    s.step(5)
    ms_step = s.step_timed(steps)
    _, ms_kernel = s.step_kernel_timed(steps)
    s.close()
    per_launch = ms_kernel / steps / 1e3
    per_step = ms_step / steps / 1e3
    algo = bytes_per_update * nodes
    return per_launch, algo / per_launch / 1e9, per_step, algo / per_step / 1e9


def traffic_from_profiles(name):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            return json.load(fh).get(name)
    except (OSError, ValueError):
        return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(seconds=15.0):
    """The FP64 CPU oracle (push-form port of the plain-mode step, OpenMP) on a slab of the same
    workload: 512x256x32 nodes with the same cylinder, walls, inlet and outlet, on all the
    threads OpenMP is given and on one thread (a shorter sample); then the same step built in
    float (oracle/liboracle_f32.so) on all threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    nx, ny, nz = 512, 256, 32
    u, D, Re = 0.05, 40.0, 300.0
    cfg = O.make_cfg(27, nx, ny, nz, tau0=3 * u * D / Re, u0=u, body=O.BODY_ZCYL, center=(128.0, 128.0, 0.0),
                     r2=(D / 2) ** 2, perturb=1e-3, force_interval=10)
    threads = O.lib().orc_num_threads()

    def sample(secs):
        st = O.PlainState(cfg)
        st.step(1)
        t0 = time.perf_counter()
        st.step(1)
        one = time.perf_counter() - t0
        n = max(2, int(secs / max(one, 1e-3)))
        t0 = time.perf_counter()
        st.step(n)
        return n, time.perf_counter() - t0

    n, dt = sample(seconds)
    O.lib().orc_set_threads(1)
    n1, dt1 = sample(seconds / 3.0)
    O.lib().orc_set_threads(threads)
    # FP32: the float build of the same step, all threads
    L32 = O.f32_lib()
    L32.orc_set_threads(threads)
    import ctypes
    h = L32.orc_plain_create(ctypes.byref(cfg), None, None, None, None)
    L32.orc_plain_step(h, 1)
    t0 = time.perf_counter()
    L32.orc_plain_step(h, 1)
    one = time.perf_counter() - t0
    n32 = max(2, int(seconds / 2.0 / max(one, 1e-3)))
    t0 = time.perf_counter()
    L32.orc_plain_step(h, n32)
    dt32 = time.perf_counter() - t0
    L32.orc_plain_destroy(h)
    return {"value": round(nx * ny * nz * n / dt / 1e6, 2), "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle, {n} steps of a 512x256x32 slab of the workload ({dt:.1f} s)",
            "one_thread": {"value": round(nx * ny * nz * n1 / dt1 / 1e6, 2), "unit": "MLUPS",
                           "sample": f"{n1} steps ({dt1:.1f} s)"},
            "fp32": {"value": round(nx * ny * nz * n32 / dt32 / 1e6, 2), "unit": "MLUPS", "cores": threads,
                     "sample": f"the oracle's step built in float, {n32} steps of the same slab ({dt32:.1f} s)"},
            "cpu_model": cpu_model(), "host_threads": os.cpu_count()}


def cpu_reference_structured(seconds=10.0):
    """The oracle's reference mode on the same 512x256x32 slab: the reference's own algorithm
    and structure (collision, six ghost fills, usrbc, stream into ft, copy back, the serial
    moment update; AoS ghost-padded layout; OpenMP on the collision and stream loops, the
    update serial as in the reference), with its boundary rules in place of the workload's.
    The nearest stand-in for the reference's CPU build, which cannot be built here."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    nx, ny, nz = 512, 256, 32
    u, D, Re = 0.05, 40.0, 300.0
    cfg = O.make_cfg(27, nx, ny, nz, tau0=3 * u * D / Re, u0=u, body=O.BODY_ZCYL, center=(128.0, 128.0, 0.0),
                     r2=(D / 2) ** 2, force_interval=10)
    st = O.RefState(cfg)
    st.step(1)
    t0 = time.perf_counter()
    st.step(1)
    one = time.perf_counter() - t0
    n = max(2, int(seconds / max(one, 1e-3)))
    t0 = time.perf_counter()
    st.step(n)
    dt = time.perf_counter() - t0
    del st
    # FP32: the float build of the same reference mode
    import ctypes
    L32 = O.f32_lib()
    h = L32.orc_ref_create(ctypes.byref(cfg), None)
    L32.orc_ref_step(h, 1)
    t0 = time.perf_counter()
    L32.orc_ref_step(h, 1)
    one = time.perf_counter() - t0
    n32 = max(2, int(seconds / 2.0 / max(one, 1e-3)))
    t0 = time.perf_counter()
    L32.orc_ref_step(h, n32)
    dt32 = time.perf_counter() - t0
    L32.orc_ref_destroy(h)
    return {"value": round(nx * ny * nz * n / dt / 1e6, 2), "unit": "MLUPS", "cores": O.lib().orc_num_threads(),
            "kind": "port", "sample": f"FP64 oracle reference mode, {n} steps of a 512x256x32 slab ({dt:.1f} s)",
            "fp32": {"value": round(nx * ny * nz * n32 / dt32 / 1e6, 2), "unit": "MLUPS",
                     "sample": f"the reference mode built in float, {n32} steps of the same slab ({dt32:.1f} s)"}}


def adjoint_ratio(pkg, steps=None):
    """Adjoint (recompute + reverse sweep) over forward time on the D3Q27 adjoint workload:
    192x96x96, FP32, J = mean Cd over steps 2000-6000, 6000 steps, checkpoint every 250
    (`steps`: a shorter window of the same workload, J over its last two thirds, for profiler
    runs; the per-launch kernels are the same)."""
    W = pkg.configs.adjoint3d_192x96x96()
    n0, n1 = W.window
    if steps:
        n0, n1 = max(1, steps // 3), steps
    s = pkg.Solver(W.cfg)
    s.set_solid_fraction(W.extra["sfrac"])
    s.initialize()
    J, ds, du, tf, ta = s.adjoint(n0, n1, W.cd_scale(), W.checkpoint)
    s.close()
    return {"workload": f"d3q27_adjoint_192x96x96 (fp32), J over steps {n0}-{n1}, checkpoint {W.checkpoint}",
            "forward_ms": round(tf, 3), "adjoint_ms": round(ta, 3), "ratio": round(ta / tf, 3), "J": J}


def dropin_loop(pkg, steps=200, warm=10):
    """The reference's own 3D driver loop through the drop-in classes (host/lbm_main on the
    3D input.in case: D3Q27 512x128x128 sphere, reference boundary mode, FP64; Collision(),
    pack/exchange/unpack, Streaming(), Update() per step, main.cpp:29-45) against one
    lbm_step(n) call on the same case: seconds per step after `warm` steps, device drained at
    both ends.  The run is capped at `steps` (--time), below the first snapshot step (inter =
    Time / 10 = 3200, System.cpp:21-22), so only the initial snapshot is written, before timing."""
    import subprocess
    import tempfile
    exe = os.path.join(PKG_DIR, "host", "lbm_main")
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "input.in")
        with open(path, "w") as fh:
            fh.write("1.0  8.0  100\n0.1  100  10\n512  128  128\n")
        r = subprocess.run([exe, "--input", path, "--time", str(steps), "--report", str(warm)], cwd=d,
                           capture_output=True, text=True, timeout=300)
        if r.returncode != 0:
            raise RuntimeError(f"lbm_main failed: {r.stderr[-500:]}")
        line = [l for l in r.stdout.splitlines() if l.startswith("loop_seconds")][-1].split()
        loop_s, n = float(line[1]), int(line[3])
    c = pkg.lbm.default_config(27)
    with pkg.Solver(c) as s:
        s.initialize()
        s.step(warm)
        s.sync()
        t0 = time.perf_counter()
        s.step(n)
        s.sync()
        step_s = time.perf_counter() - t0
    return {"workload": "3D input.in case: D3Q27 512x128x128 sphere, reference mode, fp64",
            "dropin_loop_ms_per_step": round(loop_s / n * 1e3, 5), "lbm_step_ms_per_step": round(step_s / n * 1e3, 5),
            "ratio": round(loop_s / step_s, 4), "steps": n}


def config_legs(pkg):
    """The two other forward configurations at their own sizes.  Config 0 (D2Q9 400x100 cylinder,
    FP64, drag every 10 steps): time per step over 2000 steps, several steps per launch (the
    default on small domains) against one launch per step.  Config 1 (D3Q27 256x128x128 sphere):
    MLUPS over 200 steps in each storage format, and after its 10 000 steps the FP32 and FP16
    runs' mean drag (second half) and velocity against the FP64 run."""
    import numpy as np
    out = {}
    W = pkg.configs.cylinder2d_400x100()
    n = W.cfg.nx * W.cfg.ny
    c0 = {"workload": W.name + " fp64, drag every 10 steps"}
    for tb, key in (("1", "several_steps_per_launch"), ("0", "one_launch_per_step")):
        os.environ["LBM_TEMPORAL_BLOCKING"] = tb
        with pkg.Solver(W.cfg) as s:
            s.initialize()
            s.step_timed(20)
            ms = s.step_timed(2000)
        c0[key] = {"us_per_step": round(ms / 2000 * 1e3, 3), "mlups": round(n * 2000 / ms / 1e3, 1)}
    os.environ.pop("LBM_TEMPORAL_BLOCKING", None)
    c0["speedup"] = round(c0["one_launch_per_step"]["us_per_step"] / c0["several_steps_per_launch"]["us_per_step"], 2)
    out["config0"] = c0
    log(f"config 0: {c0}")
    L = pkg.lbm
    c1 = {"workload": "d3q27_sphere_256x128x128_re200"}
    ref = None
    for name, prec in (("fp64", L.FP64), ("fp32", L.FP32), ("fp16s", L.FP16S)):
        W = pkg.configs.sphere3d_256x128x128(precision=prec)
        with pkg.Solver(W.cfg) as s:
            s.initialize()
            s.step_timed(10)
            ms = s.step_timed(200)
            s.step(W.steps - 210)
            F = s.forces()
            _, u, v, w = s.fields()
        sel = F[:, 0] > W.steps // 2
        rec = {"mlups": round(W.cfg.nx * W.cfg.ny * W.cfg.nz * 200 / ms / 1e3, 1),
               "cd_mean": float(W.cd_scale() * F[sel, 1].mean())}
        vel = np.stack([u, v, w])
        if ref is None:
            ref = (rec["cd_mean"], vel)
        else:
            rec["cd_rel_err"] = float(abs(rec["cd_mean"] - ref[0]) / abs(ref[0]))
            rec["vel_l2_rel_err"] = float(np.linalg.norm(vel - ref[1]) / np.linalg.norm(ref[1]))
        c1[name] = rec
        log(f"config 1 {name}: {rec}")
    out["config1"] = c1
    return out


def accuracy(pkg, steps, early, window_start=None):
    """Storage formats against FP64 on the benchmark workload: mean drag and RMS lift
    coefficients over steps [window_start, steps] (default: the second half), the Strouhal
    number of the lift signal band-passed to St in [0.05, 0.5] (which removes the channel's
    transverse acoustic mode, period 2 ny / c_s = 887 steps, St 0.9), and the relative L2 error
    of the velocity field after `early` steps (before the unsteady wakes drift apart in phase).
    The FP64 run is the oracle's algorithm on the GPU (bit-identical to the CPU oracle on the
    parity cases; the CPU oracle itself would need hours at this size)."""
    ws = steps // 2 if window_start is None else window_start
    import numpy as np
    L = pkg.lbm
    out, ref = {}, None
    for name, prec in (("fp64", L.FP64), ("fp32", L.FP32), ("fp16s", L.FP16S), ("bf16s", L.BF16S)):
        W = pkg.configs.zcylinder3d_512x256x256(precision=prec, force_interval=10)
        s = pkg.Solver(W.cfg).initialize()
        s.step(early)
        _, u, v, w = s.fields()
        t0 = time.perf_counter()
        done = early
        while done < steps:
            n = min(10000, steps - done)
            s.step(n)
            s.sync()
            done += n
            log(f"accuracy {name}: step {done}")
        dt = time.perf_counter() - t0
        F = s.forces()
        s.close()
        sel = F[:, 0] > ws
        cd = W.cd_scale() * F[sel, 1]
        cl = W.cd_scale() * F[sel, 2]
        rec = {"cd_mean": float(cd.mean()), "cl_rms": float(np.sqrt(np.mean((cl - cl.mean()) ** 2))),
               "strouhal": float(pkg.configs.strouhal(cl, 10.0, W.D, W.u0, band=(0.05, 0.5))),
               "seconds": round(dt, 2)}
        if ref is None:
            ref = (rec, u, v, w)
        else:
            r, ur, vr, wr = ref
            num = np.sqrt(np.sum((u - ur) ** 2) + np.sum((v - vr) ** 2) + np.sum((w - wr) ** 2))
            den = np.sqrt(np.sum(ur ** 2) + np.sum(vr ** 2) + np.sum(wr ** 2))
            rec["vel_l2_rel_err_at_early"] = float(num / den)
            for k in ("cd_mean", "cl_rms", "strouhal"):
                rec[k + "_rel_err"] = float(abs(rec[k] - r[k]) / abs(r[k]))
        del u, v, w
        out[name] = rec
        log(f"accuracy {name}: {rec}")
    out["workload"] = f"D3Q27 z-cylinder 512x256x256 Re 300, {steps} steps, statistics over steps {ws}-{steps}, " \
                      f"velocity compared at step {early}"
    return out


def copy_calibration(n_bytes=1 << 30, reps=3):
    """A plain device-to-device copy of known size (read n_bytes + write n_bytes) through the
    HIP runtime (its copy kernel): that dispatch's FETCH_SIZE / WRITE_SIZE counters give the
    scale factor from counter units to bytes."""
    import ctypes as C
    try:
        hip = C.CDLL("libamdhip64.so")
    except OSError:
        hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(C.byref(p), C.c_size_t(n_bytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(a, 1, C.c_size_t(n_bytes)) != 0:
            raise RuntimeError("hipMemset failed")
        for _ in range(reps):
            if hip.hipMemcpy(b, a, C.c_size_t(n_bytes), 3) != 0:   # hipMemcpyDeviceToDevice
                raise RuntimeError("hipMemcpy failed")
        if hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipDeviceSynchronize failed")
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    return {"bytes_read": n_bytes, "bytes_written": n_bytes, "reps": reps}


def default_configs(pkg, steps=100):
    """Throughput of the reference's own default configurations (lbm_config_default: the 3D
    input.in case, D3Q27 512x128x128 sphere, and the 2D MPI case, D2Q9 512x128 disk with drag
    every step) in reference mode (FP64, the reference's boundary rules) and in plain mode at
    each storage precision.  HIP-event time of `steps` steps after 10 warm-up steps."""
    L = pkg.lbm
    rows = []
    for Q in (27, 9):
        c = L.default_config(Q)
        for prec, bc, name in ((L.FP64, L.BC_REFERENCE, "reference fp64"), (L.FP64, L.BC_PLAIN, "plain fp64"),
                               (L.FP32, L.BC_PLAIN, "plain fp32"), (L.FP16S, L.BC_PLAIN, "plain fp16s")):
            c.precision, c.bc_mode = prec, bc
            with L.Solver(c) as s:
                s.initialize()
                s.step_timed(10)
                ms = s.step_timed(steps)
                s.sync()
            n = c.nx * c.ny * (c.nz if Q == 27 else 1)
            rows.append({"lattice": f"D{3 if Q == 27 else 2}Q{Q}", "extent": [c.nx, c.ny, c.nz if Q == 27 else 1],
                         "mode": name, "force_interval": c.force_interval, "ms_per_step": round(ms / steps, 5),
                         "mlups": round(n * steps / ms / 1e3, 1)})
            log(f"{rows[-1]['lattice']} {name}: {rows[-1]['mlups']} MLUPS")
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 1 GiB device copy (reference for profiler byte counters)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--accuracy", type=int, default=0, metavar="STEPS",
                    help="instead of timing, compare FP32/FP16/BF16 storage with FP64 over STEPS steps")
    ap.add_argument("--accuracy-early", type=int, default=2000)
    ap.add_argument("--accuracy-window", type=int, default=None, help="first step of the statistics window")
    ap.add_argument("--default-configs", action="store_true",
                    help="instead of the benchmark, time the reference's default 3D and 2D configurations")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="instead of the benchmark, time the adjoint leg alone (forward and reverse sweep)")
    ap.add_argument("--adjoint-steps", type=int, default=None, metavar="N",
                    help="with --adjoint-only: a window of N steps (J over the last two thirds) instead of 6000")
    ap.add_argument("--configs-only", action="store_true",
                    help="instead of the benchmark, run the config 0 / config 1 legs alone")
    ap.add_argument("--no-configs", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the whole domain runs on a single MI355X; "
              "multi-GPU runs are not supported by this program (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    pkg = load_package()
    L = pkg.lbm
    K, Wm = a.steps, a.warmup
    if a.accuracy:
        print(json.dumps({"accuracy": accuracy(pkg, a.accuracy, a.accuracy_early, a.accuracy_window)}), flush=True)
        return
    if a.default_configs:
        print(json.dumps({"default_configs": default_configs(pkg)}), flush=True)
        return
    if a.adjoint_only:
        print(json.dumps({"adjoint": adjoint_ratio(pkg, a.adjoint_steps)}), flush=True)
        return
    if a.configs_only:
        print(json.dumps({"configs": config_legs(pkg), "dropin_loop": dropin_loop(pkg)}), flush=True)
        return

    log(f"fp16s leg: {K} steps after {Wm} warm-up")
    s16, W16, wall16, info16 = run_leg(pkg, L.FP16S, K, Wm, 10)
    nodes = info16.nodes
    forces = s16.forces()
    s16.close()
    mlups16 = nodes * K / wall16 / 1e6
    log(f"fp16s {mlups16:.0f} MLUPS, {wall16 / K * 1e3:.4f} ms/step; kernel roofline run")
    t16, ach16, st16, sach16 = kernel_roofline(pkg, L.FP16S, K, info16.bytes_per_update, nodes)
    out = {
        "metric": METRIC, "value": round(mlups16, 1), "unit": "MLUPS", "n_gpus": 1, "steps": K, "warmup": Wm,
        "ms_per_step": round(wall16 / K * 1e3, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": W16.name, "lattice": "D3Q27", "nx": 512, "ny": 256, "nz": 256,
                   "storage": "fp16 shifted (f_i - w_i rho0)/w_i", "collision": "fp32",
                   "force_interval": 10, "body": "z-cylinder D=40", "Re": 300},
        "roofline": {"bound": "hbm", "achieved": round(ach16, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach16 / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles("fp16s_k_fast_v"),
                     "kernel": f"k_fast_v<27,SF16,{info16.nodes_per_lane}>", "kernel_ms": round(t16 * 1e3, 5),
                     "bytes_per_update": info16.bytes_per_update,
                     "step": {"kernels": "k_fast_v", "ms": round(st16 * 1e3, 5),
                              "achieved": round(sach16, 1), "frac": round(sach16 / HBM_PEAK_GBS, 4),
                              "traffic": traffic_from_profiles("fp16s")}},
        "forces_finite": bool(forces.size and (abs(forces[:, 1:]) < 1e30).all()),
    }
    if not a.no_fp32:
        log("fp32 leg")
        s32, W32, wall32, info32 = run_leg(pkg, L.FP32, K, Wm, 10)
        s32.close()
        t32, ach32, st32, sach32 = kernel_roofline(pkg, L.FP32, K, info32.bytes_per_update, nodes)
        out["fp32"] = {"value": round(nodes * K / wall32 / 1e6, 1), "unit": "MLUPS",
                       "ms_per_step": round(wall32 / K * 1e3, 5),
                       "roofline": {"bound": "hbm", "achieved": round(ach32, 1), "peak": HBM_PEAK_GBS,
                                    "unit": "GB/s", "frac": round(ach32 / HBM_PEAK_GBS, 4),
                                    "traffic": traffic_from_profiles("fp32_k_fast_v"),
                                    "kernel": f"k_fast_v<27,SF32,{info32.nodes_per_lane}>",
                                    "kernel_ms": round(t32 * 1e3, 5), "bytes_per_update": info32.bytes_per_update,
                                    "step": {"kernels": "k_fast_v", "ms": round(st32 * 1e3, 5),
                                             "achieved": round(sach32, 1), "frac": round(sach32 / HBM_PEAK_GBS, 4),
                                             "traffic": traffic_from_profiles("fp32")}}}
    if not a.no_adjoint:
        log("adjoint leg")
        out["adjoint"] = adjoint_ratio(pkg)
    if not a.no_configs:
        log("config 0 / config 1 legs")
        out["configs"] = config_legs(pkg)
        log("drop-in loop")
        out["dropin_loop"] = dropin_loop(pkg)
    if not a.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(a.cpu_seconds)
        log("cpu reference-structured")
        out["cpu_reference_structured"] = cpu_reference_structured(a.cpu_seconds * 2.0 / 3.0)
    if a.calibrate_copy:
        log("copy calibration")
        out["copy_calibration"] = copy_calibration()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
