"""CPU emulation of the persistent kernels' work queues (replacing the
reference's NXTVAL task counter, ccsd_t_fused_driver.hpp:169-172,381,456).

The library builds the queue layout a run would launch with
(xt_queue_layout, host only); tests/host/queue_emu.cpp replays the kernels'
claim loop over it with the kernels' own ticket decode (xt_queue.h), under
the address and undefined-behaviour sanitizers, in seeded random claim
orders, and checks that every box of every selected task is claimed exactly
once and that every workgroup leaves the loop.  Covered: FP64 sub-tiles and
mixed boxes, every schedule, several ticket-group shapes, task subsets and
grid sizes from 1 to 512 workgroups."""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("emu") / "queue_emu")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all", os.path.join(HERE, "host", "queue_emu.cpp"),
                    "-o", exe], check=True)
    return exe


def run_emu(exe, path, lay, sel, grid, seed):
    with open(path, "wb") as f:
        np.array([len(lay["nbox"]), len(lay["qtask"])], dtype=np.int64).tofile(f)
        lay["qbeg"].astype(np.int32).tofile(f)
        lay["qtask"].astype(np.int32).tofile(f)
        lay["qtick"].astype(np.int64).tofile(f)
        lay["qtotal"].astype(np.int64).tofile(f)
        lay["nbox"].astype(np.int32).tofile(f)
        lay["glog"].astype(np.int32).tofile(f)
        sel.astype(np.int32).tofile(f)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([exe, path, str(grid), str(seed)], capture_output=True, text=True, env=env)
    assert r.returncode == 0 and "queue_emu: ok" in r.stdout, r.stdout + r.stderr
    return r.stdout


TILINGS = {
    "config0": (8, 48, [8], [16, 16, 16]),
    "uneven_small": (7, 22, [3, 4], [5, 9, 8]),
    "wide_tiles": (7, 53, [6, 1], [50, 3]),
}
SHAPES = [[1, 1, 1, 1, 1, 1], [2, 2, 2, 0, 0, 0], [0, 0, 0, 2, 2, 2], [6, 0, 0, 0, 0, 0],
          [0, 0, 0, 0, 0, 6], [0, 1, 1, 1, 1, 2]]


def make_ctx(xt, name):
    ctx = xt.CCSDTriples(-1)
    if name == "uneven_o20_v120":
        ctx.set_tiling(xt.Tiling.uneven(20, 120, (6, 20), (16, 48), seed=7))
    else:
        O_, V_, occ, vir = TILINGS[name]
        ctx.set_tiling(xt.Tiling(O_, V_, occ, vir))
    return ctx


@pytest.mark.parametrize("name", list(TILINGS) + ["uneven_o20_v120"])
@pytest.mark.parametrize("prec", [0, 1])
def test_every_box_claimed_once(xt, emu, tmp_path, name, prec):
    ctx = make_ctx(xt, name)
    n = ctx.num_tasks()
    rng = np.random.default_rng(5)
    ctx.set_task_cost(rng.uniform(0.5, 2.0, n) * ctx.flops()[1])  # installs learned costs
    subsets = [None, np.sort(rng.choice(n, max(1, n // 3), replace=False))[::-1]]
    shapes = SHAPES if name != "uneven_o20_v120" else SHAPES[:2]
    k = 0
    for shape in shapes:
        ctx.set_group_shape(prec, shape)
        for sched in (xt.SCHED_DRIVER, xt.SCHED_ANALYTIC, xt.SCHED_LEARNED):
            for which in subsets:
                lay = ctx.queue_layout(which, precision=prec, schedule=sched)
                assert list(lay["glog"]) == shape
                sel = np.zeros(n, dtype=np.int32)
                sel[np.arange(n) if which is None else which] = 1
                for grid in (1, 7, 512 if prec == 0 else 256):
                    out = run_emu(emu, str(tmp_path / "lay.bin"), lay, sel, grid, 100 + k)
                    k += 1
    print(out)
    ctx.close()


def test_schedules_order_but_never_change_the_task_set(xt):
    """Driver order deals tasks round-robin in the reference's loop order;
    the cost schedules place the largest first, each queue in descending
    cost, every queue's load within one task of the others (LPT)."""
    ctx = make_ctx(xt, "uneven_o20_v120")
    n = ctx.num_tasks()
    _, per_task, _ = ctx.flops()
    d = ctx.queue_layout(schedule=xt.SCHED_DRIVER)
    for q in range(8):
        ent = d["qtask"][d["qbeg"][q]:d["qbeg"][q + 1]]
        assert list(ent) == list(range(q, n, 8))
    a = ctx.queue_layout(schedule=xt.SCHED_ANALYTIC)
    assert sorted(a["qtask"]) == list(range(n))
    loads = []
    for q in range(8):
        ent = a["qtask"][a["qbeg"][q]:a["qbeg"][q + 1]]
        c = per_task[ent]
        assert np.all(np.diff(c) <= 0)
        loads.append(c.sum())
    assert max(loads) - min(loads) <= per_task.max()
    ctx.close()
