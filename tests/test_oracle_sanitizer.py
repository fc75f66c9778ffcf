"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer: `make asan_check` builds
a driver (oracle/asan_check.c) that calls every oracle entry point on small cases -- reference
mode (3D, the 2D MPI and single-node variants, staged and individual kernels), plain mode
(sphere, mask, solid fraction, injected perturbation and inlet profile, periodic axes) and the
adjoint -- and any sanitizer report aborts it with a non-zero status.  CPU only."""
import os
import subprocess

ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


def test_oracle_runs_clean_under_asan_and_ubsan():
    subprocess.run(["make", "-s", "-B", "-C", ORACLE_DIR, "asan_check"], check=True)
    env = dict(os.environ, OMP_NUM_THREADS="2", ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(ORACLE_DIR, "asan_check")], env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "asan_check: ok" in r.stdout
