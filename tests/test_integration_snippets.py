"""The C++ caller code INTEGRATION.md shows a maintainer, compiled against
include/exachem_t.h and linked against the built library.

Every ```cpp block of INTEGRATION.md is taken as written.  The first one (the
per-task launcher stub, a template at namespace scope) is compiled at file
scope and instantiated for double; each later one is the body of a function.
A prelude stands in for the names the blocks take from ExaChem's driver:
TAMM's stream pair, event pointer and tamm_terminate, and the driver's
variables (tile indices, packed buffers, sizes, inputs).  The program is
linked with -lexachem_t, so every entry point the blocks call must be
exported, and it runs (calling nothing) to show the link is complete."""
import os
import re
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

PRELUDE = r"""
#include <string>
#include <type_traits>
#include <utility>
#include <vector>
#include <cstdio>
#include <cstdlib>
#include "exachem_t.h"

// stand-ins for the names ExaChem's driver provides around the blocks
using gpuStream_t = std::pair<void*, int>;   // TAMM keeps the HIP stream in .first
using event_ptr_t = void**;                  // points at a hipEvent_t
[[noreturn]] static void tamm_terminate(const std::string& m) { std::fprintf(stderr, "%s\n", m.c_str()); std::abort(); }
static xt_ctx*     xt = nullptr;
static gpuStream_t stream{nullptr, 0};
static size_t      t_h1b, t_h2b, t_h3b, t_p4b, t_p5b, t_p6b;
static double     *df_dev_d1_t2_all, *df_dev_d1_v2_all, *df_dev_d2_t2_all, *df_dev_d2_v2_all,
                  *df_dev_s1_t1_all, *df_dev_s1_v2_all;
static int         df_simple_s1_size[6], df_simple_s1_exec[9], df_simple_d1_size[7 * 50],
                   df_simple_d1_exec[9 * 50], df_simple_d2_size[7 * 140], df_simple_d2_exec[9 * 140];
static int         df_num_s1_enabled, df_num_d1_enabled, df_num_d2_enabled;
static double     *eps_occ, *eps_vir, *t1, *t2, *vvoo, *ovoo, *vvvo, *t1_aa, *t2_abab, *chol_spatial;
static long        nchol;
static int         O, V, n_occ_tiles, n_vir_tiles, occ_tiles[8], vir_tiles[16], is_restricted;
static double      energy1, energy2;
static double*     partials;
static long        n_tasks;
static xt_opts     o;
static double      e45[2];
static std::vector<long> tasks_without_partials() { return {}; }
"""


def blocks():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    return re.findall(r"```cpp\n(.*?)```", text, re.S)


def test_integration_blocks_found():
    b = blocks()
    assert len(b) >= 6, len(b)
    assert "xt_fully_fused_ccsd_t_gpu" in b[0]
    calls = set(re.findall(r"\b(xt_\w+)\s*\(", "".join(b)))
    import exachem_amd as X
    assert calls <= set(X.header_symbols()), calls - set(X.header_symbols())


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_integration_blocks_compile_and_link(xt, tmp_path):
    b = blocks()
    src = [PRELUDE, b[0],
           "template void fully_fused_ccsd_t_gpu_mi355x<double>(gpuStream_t&, size_t, size_t, "
           "size_t, size_t, size_t, size_t, size_t, double*, double*, double*, double*, double*, "
           "double*, int*, int*, int*, int*, int*, int*, size_t, size_t, size_t, size_t, size_t, "
           "size_t, size_t, size_t, double*, double*, double*, double*, double*, double*, double*, "
           "event_ptr_t);"]
    for k, body in enumerate(b[1:], 1):
        src.append(f"void snippet_{k}() {{\n{body}\n}}")  # external: emitted and linked
    src.append("int main() { std::puts(\"linked\"); return 0; }")
    cpp = tmp_path / "callers.cpp"
    cpp.write_text("\n".join(src))
    exe = str(tmp_path / "callers")
    libdir = os.path.join(ROOT, "exachem_amd", "_lib")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-Wno-unused-function",
                        "-Wno-unused-variable", "-I", os.path.join(ROOT, "include"), str(cpp),
                        "-o", exe, "-L", libdir, "-lexachem_t", f"-Wl,-rpath,{libdir}",
                        "-Wl,-rpath,/opt/rocm/lib", "-L/opt/rocm/lib"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ)
    env.pop("LD_PRELOAD", None)
    run = subprocess.run([exe], capture_output=True, text=True, env=env)
    assert run.returncode == 0 and run.stdout.strip() == "linked", run.stdout + run.stderr
