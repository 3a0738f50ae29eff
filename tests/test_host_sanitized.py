"""The library's host code under AddressSanitizer and UBSan, on the CPU.

tests/cpu_hip/ builds the host side of every translation unit of
libchemsim_hip.so (kernels left out) against a CPU model of the HIP runtime
whose device memory is exactly-sized heap memory and whose asynchronous copies
run only when a synchronisation forces them.  Two programs run on it: a driver
of every C-ABI entry point (storages, boundary modes, operators, render,
recording with and without a stop, batches, the network) and the C++ drop-in
layer's own test driver (tests/cpp/lbm_host_test.cpp: State, render and
record of a State, population edits between steps, a throwing frame function).

What this pins: no host or device buffer is touched outside its bounds or
after it is freed -- including the late writes of a copy whose host buffer
died before the copy was waited for, and calls into the library made from
static destructors at process exit (a thread-local error message destroyed
before them was once written to after its buffer was freed, which corrupts
glibc's free lists)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpu_hip")
OUT = os.path.join(SRC, "build")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


@pytest.fixture(scope="module")
def built():
    jobs = str(min(8, os.cpu_count() or 1))
    r = subprocess.run(["make", f"-j{jobs}", "-C", SRC, f"OUT={OUT}"], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    return OUT


def test_every_entry_point_sanitized(built):
    r = subprocess.run([os.path.join(built, "asan_driver")], capture_output=True, text=True, timeout=300,
                       env=ENV)
    assert r.returncode == 0 and "sanitized host run: ok" in r.stdout, r.stdout[-4000:] + r.stderr[-8000:]


def test_host_layer_sanitized(built, tmp_path):
    r = subprocess.run([os.path.join(built, "lbm_host_test"), str(tmp_path)], capture_output=True, text=True,
                       timeout=300, env=ENV)
    assert r.returncode == 0 and "host layer run: ok" in r.stdout, r.stdout[-4000:] + r.stderr[-8000:]
    for name in ("render.u8", "record.u8", "edit.f32", "stop.txt"):
        assert (tmp_path / name).stat().st_size > 0
