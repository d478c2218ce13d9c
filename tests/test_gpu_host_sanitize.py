"""The plan builder and host entry points under host ASan + UBSan.

tests/host_sanitize/plan_sanitize is libcint_amd/csrc/api.hip compiled with
AddressSanitizer and UndefinedBehaviorSanitizer on its host code, linked with
a driver that exercises plan construction (threaded counting sort, class
buckets, pair table, scattered offsets, FP32 flags, every kind, empty and
invalid lists), execution, the J/K and gradient digests, the per-call
functions and the batches (see the driver's header).  Built by
``make -C tests/host_sanitize`` after the library; any sanitizer finding
aborts the driver.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "host_sanitize", "_build", "plan_sanitize")
SUPP = os.path.join(HERE, "host_sanitize", "lsan.supp")


@pytest.mark.gpu
def test_plan_builder_under_host_sanitizers():
    assert os.path.exists(EXE), "build it with: make -C tests/host_sanitize"
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0:halt_on_error=1"
    env["LSAN_OPTIONS"] = f"suppressions={SUPP}:print_suppressions=0"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    print("running", EXE, flush=True)
    r = subprocess.run([EXE], env=env, capture_output=True, text=True, timeout=600)
    print(r.stdout[-4000:])
    print(r.stderr[-8000:])
    assert r.returncode == 0, f"exit status {r.returncode}"
    assert "plan_sanitize ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
