"""The host-side C++ and C under the sanitizers (CPU only): the threaded
topology and near-field pattern builders and the multi-RHS pass planner
(host_geometry.cpp: every tile in one pass, each output's summation order
the same for every partial limit) and the CPU
oracle (oracle.c) built with AddressSanitizer + UndefinedBehaviorSanitizer,
and the builders with ThreadSanitizer, running tests/sanitize/host_check.cpp
on valid and invalid meshes.  A sanitizer report or a failed check fails the
test.  (The solve server's scheduling needs a device and is covered by the
GPU tests.)"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "easybem_amd", "csrc")
SRC = os.path.join(ROOT, "tests", "sanitize", "host_check.cpp")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


def _build_and_run(tmp_path, name, flags, with_oracle):
    exe = tmp_path / name
    objs = []
    if with_oracle:
        o = tmp_path / "oracle.o"
        subprocess.run(["gcc", "-O1", "-g", "-fopenmp", *flags, "-c",
                        os.path.join(ROOT, "oracle", "oracle.c"), "-o", str(o)], check=True)
        objs.append(str(o))
    cmd = ["g++", "-std=c++17", "-O1", "-g", *flags, "-I" + CSRC,
           "-I" + os.path.join(ROOT, "oracle"), SRC, os.path.join(CSRC, "host_geometry.cpp"),
           *objs, "-o", str(exe), "-lpthread"]
    if with_oracle:
        cmd[1:1] = ["-DWITH_ORACLE", "-fopenmp"]
    subprocess.run(cmd, check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1", TSAN_OPTIONS="halt_on_error=1",
               OMP_NUM_THREADS="4")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=600)
    out = r.stdout + r.stderr
    print(out[-3000:])
    return r.returncode, out


def test_host_code_under_address_and_undefined_sanitizers(tmp_path):
    rc, out = _build_and_run(tmp_path, "host_check_asan",
                             ["-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                              "-fno-omit-frame-pointer"], with_oracle=True)
    assert rc == 0 and "0 failures" in out
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out
    assert "ERROR: LeakSanitizer" not in out


def test_threaded_builders_under_thread_sanitizer(tmp_path):
    rc, out = _build_and_run(tmp_path, "host_check_tsan", ["-fsanitize=thread"],
                             with_oracle=False)
    assert rc == 0 and "0 failures" in out
    assert "WARNING: ThreadSanitizer" not in out
