"""libmcint's host plan builder under host ASan + UBSan (+ LeakSanitizer).

tests/host_sanitize/ builds libcint_amd/csrc/api.hip with AddressSanitizer
and UndefinedBehaviorSanitizer on its host code (``make -C
tests/host_sanitize``, run by the package build) and links it with two
drivers:

- plan_host (CPU, no GPU): the host stage of plan construction over a
  3.1M-tuple list and every integral code, with its task counts, output
  sizes, bucket sums, determinism and refusals checked;
- plan_sanitize (GPU): every host entry point end to end, execution and the
  J/K and gradient digests included.

Any sanitizer finding aborts the driver with a nonzero status.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SAN_DIR = os.path.join(HERE, "host_sanitize")
SUPP = os.path.join(SAN_DIR, "lsan.supp")


def _run(name: str, ok_line: str):
    exe = os.path.join(SAN_DIR, "_build", name)
    if not os.path.exists(exe):
        import libcint_amd
        if not os.path.exists(libcint_amd.LIB_PATH):
            from libcint_amd import build
            build.build()
        subprocess.run(["make", "-s", "-C", SAN_DIR], check=True)
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:halt_on_error=1"
    env["LSAN_OPTIONS"] = f"suppressions={SUPP}:print_suppressions=0"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    print("running", exe, flush=True)
    r = subprocess.run([exe], env=env, capture_output=True, text=True, timeout=600)
    print(r.stdout[-4000:])
    print(r.stderr[-8000:])
    assert r.returncode == 0, f"exit status {r.returncode}"
    assert ok_line in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
    assert "LeakSanitizer" not in r.stderr


def test_plan_host_stage_under_host_sanitizers():
    _run("plan_host", "plan_host ok")


@pytest.mark.gpu
def test_plan_builder_under_host_sanitizers():
    _run("plan_sanitize", "plan_sanitize ok")
