"""CPU tests under the address and undefined-behaviour sanitizers.

* The oracle (oracle/ccsd_t_oracle.cpp) built with g++ -fsanitize=address,
  undefined runs a selection of its own tests: the reference's known
  answers, brute force and the spatial formula, the special cases, and the
  planted mistakes.
* The library's host code (planner, task list and flop counts, queue
  layouts, gather planning, memory guard, input checks) built with the
  sanitizers on the host side and linked with the regular device objects
  runs the host-only tests of the product (host-only contexts, no GPU).

Each runs in a child process with the sanitizer runtime preloaded; a report
ends that process with an error, which fails the test."""
import glob
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "exachem_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"


def _env(preload):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:protect_shadow_gap=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="4")
    env["LD_PRELOAD"] = preload
    return env


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_oracle_under_sanitizers(tmp_path):
    so = str(tmp_path / "libccsd_t_oracle.so")
    subprocess.run(["g++", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-fno-omit-frame-pointer", "-fopenmp", "-fPIC", "-std=c++17", "-shared", "-o",
                    so, os.path.join(ROOT, "oracle", "ccsd_t_oracle.cpp")], check=True)
    asan = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True,
                          text=True).stdout.strip()
    env = _env(asan)
    env["XT_ORACLE_LIB"] = so
    sel = ("total_num_ops_known_answers or bruteforce_and_spatial or special_cases or "
           "planted_mistakes or on_the_fly")
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(HERE, "test_oracle.py"), "-q",
                        "-x", "-p", "no:cacheprovider", "-k", sel], cwd=ROOT, env=env,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert " passed" in r.stdout


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="needs hipcc")
def test_library_host_code_under_sanitizers(xt, tmp_path):
    objs = os.path.join(ROOT, "exachem_amd", "_lib", "obj")
    dev = [os.path.join(objs, f) for f in ("xt_kernels.o", "xt_mixed.o", "xt_mlp.o",
                                           "xt_gather_dev.o")]
    if not all(os.path.exists(f) for f in dev):
        pytest.skip("regular build objects missing")
    # host translation units with the sanitizers on the host side only
    san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-omit-frame-pointer"]
    host = []
    for f in ("xt_host", "xt_cost", "xt_gather", "xt_chol"):
        o = str(tmp_path / f"{f}.o")
        subprocess.run([HIPCC, "-O1", "-g", "-std=c++17", "-fPIC", "--offload-host-only", *san,
                        "-c", "-o", o, os.path.join(CSRC, f + ".cpp")], check=True, cwd=CSRC,
                       capture_output=True)
        host.append(o)
    so = str(tmp_path / "libexachem_t.so")
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *san, "-shared-libsan",
                    "-o", so, *host, *dev, "-lpthread", "-L/opt/rocm/lib", "-lrocblas",
                    "-Wl,-rpath,/opt/rocm/lib"], check=True, capture_output=True)
    rt = glob.glob("/opt/rocm/lib/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")
    if not rt:
        pytest.skip("no clang sanitizer runtime")
    env = _env(rt[0])
    env["LD_LIBRARY_PATH"] = os.path.dirname(rt[0])
    code = ("import sys, pytest, exachem_amd as x\n"
            f"x.LIB_PATH = {so!r}\n"
            "assert x.lib()._name == x.LIB_PATH\n"
            "sys.exit(pytest.main(['tests/test_plan.py', 'tests/test_host_safety.py', '-q', '-x',"
            " '-p', 'no:cacheprovider', '-k', 'not fixed_point']))\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert " passed" in r.stdout
