"""Benchmark of the closed-shell CCSD(T) triples energy on one MI355X.

One step = one full (T) energy over every task of the workload: host queue
build, one graph launch (fill kernels, persistent kernel, reductions) and the
copy of E[4]/E[5] back; inputs are resident in HBM before timing starts and
each step returns only after the stream has drained, so the wall clock around
it brackets the device work.  Default workload: O=48, V=384, seeded uneven
tiles (occupied U[6,20], virtual U[16,48]), synthetic inputs of the stated
distributions (integrals N(0, 0.05^2), T1 N(0, 1e-6), MP2-like T2), mixed
precision.

Prints ONE JSON line.  value = time-to-E(T) in seconds (lower is better).
A wall-clock budget (--budget-s, counted from process start) caps the number
of warm-up and timed steps: at O=48, V=384 one step is about two minutes, so
the timed steps are as many of those requested as fit after the bounded legs
below, at least one; the line reports steps and warm-up requested beside those
run.  The calibration launches (autotuner, cost model, comparison sample) run
the same kernel and graph before the timed steps.
roofline: the reference's analytic flop count per launch over the persistent
kernel's HIP-event time.  Beside it:
  * fp64_compare: the plain FP64 kernel and the measured precision on the same
    seeded task sample (or on the whole workload with --fp64-full): time ratio
    and |dE4|, |dE5| of the measured precision against FP64;
  * schedule_ab: the measured precision on the same comparison sample under
    the reference's driver order, the analytic flop-count schedule and the
    learned schedule (same kernel, same group shape);
  * oracle_check: GPU FP64 and the measured precision, task by task, against
    the oracle's energies committed in tests/golden for this workload's seeded
    inputs (written on the CPU by tests/golden/make_golden.py), and the same at
    a physical energy scale (integrals scaled so that |E(T)| is below 1 Eh);
  * cpu_baseline: the CPU oracle (a C++ restatement of the reference's CPU
    fused (T) path) timed on a seeded, size-stratified task sample (two
    tasks per flop quartile, smallest first) in a host thread that runs
    beside the GPU work from the upload to the end of the timed steps, on
    all host cores but one, extrapolated by the flop count; its energies on
    those tasks are checked against the GPU too (cpu_sample_check).
"""
import argparse
import ctypes as C
import json
import os
import sys
import threading
import time

import numpy as np

T_START = time.time()
ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"fp64": 78.6, "mixed": 2500.0}   # MI355X dense matrix peaks (FP64, FP16)
SEED = 20241202

CONFIGS = {
    # name: (O, V, tiling spec)
    "c0": (8, 48, ("uniform", 8, 16)),
    "c1": (40, 300, ("uniform", 16, 32)),
    "c3": (48, 384, ("uneven", (6, 20), (16, 48))),
    "c4": (64, 512, ("uneven", (6, 20), (16, 48))),
}


def make_tiling(xt, name):
    O_, V_, spec = CONFIGS[name]
    if spec[0] == "uniform":
        return xt.Tiling.uniform(O_, V_, spec[1], spec[2])
    return xt.Tiling.uneven(O_, V_, spec[1], spec[2], seed=SEED)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def elapsed():
    return time.time() - T_START


def golden_case(name):
    path = os.path.join(ROOT, "tests", "golden", name)
    return json.load(open(path)) if os.path.exists(path) else None


def fixture_check(xt, ctx, case, precisions):
    """Every task of a committed oracle fixture run on the GPU at each
    precision; max abs/rel differences of E4, E5 against the oracle, and
    whether FP64 is within 1e-10 Eh + 2e-14 |E| on every task."""
    which = np.asarray(case["which"], dtype=np.int64)
    ref = np.asarray(case["task_partials"])
    out = {"tasks": int(len(which)), "roles": case.get("roles"), "sigma": case["sigma"],
           "reference": "oracle (CPU restatement of the reference path), committed fixture "
                        f"tests/golden/{case['file']}",
           "E4_sum_oracle": float(ref[:, 0].sum()), "E5_sum_oracle": float(ref[:, 1].sum())}
    for name, p in precisions:
        g = ctx.run_tasks(which, precision=p)["task_partials"]
        d = np.abs(g - ref)
        r = {"max_abs_dE4": float(d[:, 0].max()), "max_abs_dE5": float(d[:, 1].max()),
             "max_rel_dE4": float(np.max(d[:, 0] / np.abs(ref[:, 0]))),
             "max_rel_dE5": float(np.max(d[:, 1] / np.abs(ref[:, 1]))),
             "abs_dE4_sum": float(abs(g[:, 0].sum() - ref[:, 0].sum())),
             "abs_dE5_sum": float(abs(g[:, 1].sum() - ref[:, 1].sum()))}
        if p == xt.PREC_FP64:
            r["within_1e-10_plus_2e-14_rel"] = bool(np.all(d <= 1e-10 + 2e-14 * np.abs(ref)))
        out[name] = r
    return out


def gather_leg(xt, ctx, tiling, s, n_tasks, cpu_tasks=4):
    """The per-task path's gather (ccsd_t_data_{s1,d1,d2}_new) on the device
    (xt_gather_task) for a seeded sample of the workload's tasks, timed with
    HIP events on its stream; beside it the oracle's restatement of the host
    gathers from the same spatial inputs on one host core, into buffers
    allocated once as the driver does.  Algorithmic bytes: 8 written + at
    least 8 read per packed double."""
    from exachem_amd.hipmem import DevBuf, Event, Stream
    tasks, _ = ctx.task_list()
    rng = np.random.default_rng(7)
    pick = np.sort(rng.choice(len(tasks), min(n_tasks, len(tasks)), replace=False))
    sz = ctx.packed_sizes()
    names = ("d1_t2", "d1_v2", "d2_t2", "d2_v2", "s1_t1", "s1_v2")
    dev = {n: DevBuf(8 * sz[n]) for n in names}
    ptrs = {n: dev[n].ptr.value for n in names}
    st, e0, e1 = Stream(), Event(), Event()
    ctx.gather_task(tasks[pick[0]], ptrs, st.h)  # warm-up (first launch)
    st.sync()
    gpu_ms_task, elems = np.zeros(len(pick)), 0
    for i, k in enumerate(pick):
        e0.record(st.h)
        ctx.gather_task(tasks[k], ptrs, st.h)
        e1.record(st.h)
        gpu_ms_task[i] = e1.elapsed_ms(e0)
        elems += ctx.gather_last_elems()
    gpu_ms = float(gpu_ms_task.sum())
    for b in dev.values():
        b.free()
    for x in (st, e0, e1):
        x.destroy()
    out = {"tasks": int(len(pick)), "doubles_written": int(elems), "gpu_ms": gpu_ms,
           "gpu_ms_per_task": gpu_ms / len(pick),
           "roofline": {"bound": "hbm", "achieved": 16 * elems / gpu_ms / 1e6, "peak": 8000.0,
                        "unit": "GB/s", "frac": 16 * elems / gpu_ms / 1e6 / 8000.0,
                        "bytes_per_double": "8 written + 8 read (lower bound)"}}
    if cpu_tasks > 0:
        from oracle import oracle as O
        O.build()
        occ, vir = tiling.occ, tiling.vir
        g = O.SpatialGather(O.Tiling(occ, occ, vir, vir), s)
        sub = pick[:cpu_tasks]
        t0, n_cpu = time.perf_counter(), 0
        for k in sub:
            n_cpu += g(tasks[k])
        cpu_ms = (time.perf_counter() - t0) * 1e3
        gpu_sub = gpu_ms_task[:len(sub)].sum()
        out["cpu_tasks"] = int(len(sub))
        out["cpu_ms"] = cpu_ms
        out["cpu_threads"] = 1
        out["cpu_doubles_written"] = int(n_cpu)
        out["speedup_vs_cpu"] = cpu_ms / gpu_sub
    return out


def dropin_leg(xt, ctx, tiling, s, per_task, n_tasks):
    """The per-task drop-in path end to end on the device for a seeded sample
    of the workload's tasks: xt_gather_task writes the packed buffers
    (ccsd_t_data_{s1,d1,d2}_new) and xt_fully_fused_ccsd_t_gpu runs the
    reference launcher's contract on them (ccsd_t_all_fused.hpp:190-267),
    each timed with HIP events on their stream.  Beside it, the persistent
    FP64 kernel (xt_run_tasks) on the same tasks.  Each task's energy is the
    factor times the sum of its block partials, as hostEnergyReduce does."""
    from exachem_amd.hipmem import DevBuf, Event, Stream
    tasks, fac = ctx.task_list()
    rng = np.random.default_rng(7)
    pick = np.sort(rng.choice(len(tasks), min(n_tasks, len(tasks)), replace=False))
    sz = ctx.packed_sizes()
    names = ("d1_t2", "d1_v2", "d2_t2", "d2_v2", "s1_t1", "s1_v2")
    dev = {n: DevBuf(8 * sz[n]) for n in names}
    ptrs = {n: dev[n].ptr.value for n in names}
    max_dims = {n: sz[n] // (9 if n.startswith("s1") else
                             (sz["max_d1"] if n.startswith("d1") else sz["max_d2"]))
                for n in names}
    tiles = list(tiling.occ) * 2 + list(tiling.vir) * 2
    offs = np.concatenate([[0], np.cumsum(tiles)]).astype(np.int64)
    evl = DevBuf.from_array(np.concatenate([s["eo"], s["eo"], s["ev"], s["ev"]]))
    noab, nvab = 2 * len(tiling.occ), 2 * len(tiling.vir)
    st, e0, e1, e2 = Stream(), Event(), Event(), Event()
    gather_ms, kernel_ms, e = 0.0, 0.0, np.zeros((len(pick), 2))
    for i, k in enumerate([pick[0]] + list(pick)):  # first pass: warm-up
        dims = [int(tiles[b]) for b in tasks[k]]
        nb = int(np.prod([(d + 3) // 4 for d in dims]))
        part = DevBuf(16 * nb)
        ev = [C.c_void_p(evl.ptr.value + 8 * int(offs[b])) for b in tasks[k]]
        e0.record(st.h)
        m, _ = ctx.gather_task(tasks[k], ptrs, st.h)
        e1.record(st.h)
        xt.fully_fused_ccsd_t_gpu(st.h, nb, dims, [dev[n].ptr for n in names], m, noab, nvab,
                                  max_dims, ev, part.ptr)
        e2.record(st.h)
        p = part.to_array(2 * nb)  # synchronous: the stream's work is done
        part.free()
        if i > 0:
            gather_ms += e1.elapsed_ms(e0)
            kernel_ms += e2.elapsed_ms(e1)
            e[i - 1] = fac[k] * np.sum(p[:nb]), fac[k] * np.sum(p[nb:])
    for b in list(dev.values()) + [evl]:
        b.free()
    for x in (st, e0, e1, e2):
        x.destroy()
    r = ctx.run_tasks(pick, precision=xt.PREC_FP64)
    flops = float(per_task[pick].sum())
    d = np.abs(e - r["task_partials"]) / np.maximum(np.abs(r["task_partials"][:, :1]), 1.0)
    tf = flops / kernel_ms / 1e9
    return {"tasks": int(len(pick)), "flops": flops, "gather_ms": gather_ms,
            "packed_kernel_ms": kernel_ms, "packed_tflops": tf,
            "persistent_fp64_ms": r["kernel_ms"], "persistent_fp64_tflops":
            flops / r["kernel_ms"] / 1e9,
            "max_rel_dE_task_vs_persistent": float(d.max()),
            "roofline": {"bound": "mfma", "achieved": tf, "peak": 78.6, "unit": "TFLOP/s",
                         "frac": tf / 78.6, "kernel": "k_packed_fp64"}}


def expand_leg(ctx, tiling):
    """Dense spin-orbital T1/T2 of the workload (cd_ccsd_cs_ann.cpp:932-976)
    written by xt_expand_amplitudes, timed with HIP events on its stream;
    algorithmic bytes 8 written per element (the spatial amplitudes read are
    (2*2)^2 times fewer and stay in cache)."""
    from exachem_amd.hipmem import DevBuf, Event, Stream
    O_, V_ = tiling.O, tiling.V
    n1, n2 = 4 * V_ * O_, 16 * (V_ * O_) ** 2
    b1, b2 = DevBuf(8 * n1), DevBuf(8 * n2)
    st, e0, e1 = Stream(), Event(), Event()
    ctx.expand_amplitudes(b1.ptr.value, b2.ptr.value, st.h)  # warm-up
    st.sync()
    e0.record(st.h)
    ctx.expand_amplitudes(b1.ptr.value, b2.ptr.value, st.h)
    e1.record(st.h)
    ms = e1.elapsed_ms(e0)
    for x in (b1, b2, st, e0, e1):
        (x.free if hasattr(x, "free") else x.destroy)()
    return {"t1_doubles": n1, "t2_doubles": n2, "gpu_ms": ms,
            "roofline": {"bound": "hbm", "achieved": 8.0 * (n1 + n2) / ms / 1e6, "peak": 8000.0,
                         "unit": "GB/s", "frac": 8.0 * (n1 + n2) / ms / 1e6 / 8000.0}}


def cholesky_leg(xt, tiling, kfac, cpu=True):
    """V2 blocks from Cholesky vectors (v2tensors.cpp:52-98 in spatial
    closed-shell form): seeded vectors L [O+V][O+V][K], K = kfac*(O+V),
    symmetric in the orbital pair; xt_upload_cholesky timed end to end (host
    to device copies of L, T1, T2, the DGEMMs over K and the operand layouts),
    against numpy forming the same integrals on the host cores for a sample of
    the V^3 O block, extrapolated by flops.  T2 is MP2-like from (ai|bj)."""
    O_, V_ = tiling.O, tiling.V
    nmo, K = O_ + V_, kfac * (O_ + V_)
    rng = np.random.default_rng(11)
    L = rng.standard_normal((nmo, nmo, K))
    L += L.transpose(1, 0, 2)
    L *= 0.5 * (0.05 / np.sqrt(K)) ** 0.5
    eo = np.sort(rng.uniform(-2.0, -0.3, O_))
    ev = np.sort(rng.uniform(0.1, 3.0, V_))
    lvo = np.ascontiguousarray(L[O_:, :O_]).reshape(V_ * O_, K)
    t0 = time.perf_counter()
    vvoo = (lvo @ lvo.T).reshape(V_, O_, V_, O_).transpose(0, 2, 1, 3)
    host_vvoo_s = time.perf_counter() - t0
    D = (eo[None, None, :, None] + eo[None, None, None, :]
         - ev[:, None, None, None] - ev[None, :, None, None])
    t2 = np.ascontiguousarray(vvoo / D)
    del vvoo, D
    t1 = 1e-3 * rng.standard_normal((V_, O_))
    ctx = xt.CCSDTriples(0)
    ctx.set_tiling(tiling)
    t0 = time.perf_counter()
    ctx.upload_cholesky(eo, ev, t1, t2, L)
    gpu_s = time.perf_counter() - t0
    flops = 2.0 * K * (O_ * V_ * O_ * O_ + (V_ * O_) ** 2 + V_ * V_ * V_ * O_)
    out = {"O": O_, "V": V_, "nchol": K, "chol_bytes": int(L.nbytes), "flops": flops,
           "gpu_upload_s": gpu_s, "gpu_tflops_incl_copies": flops / gpu_s / 1e12,
           "vvvo_bytes_avoided": int(8 * V_ ** 3 * O_)}
    ctx.close()
    if cpu:
        # (ab|ci) for 8 values of a, the rows of the V^3 O block numpy forms fastest
        lvv = L[O_:, O_:]
        t0 = time.perf_counter()
        for a in range(8):
            _ = lvv[a].reshape(V_, K) @ lvo.T
        s_cpu = time.perf_counter() - t0
        f_cpu = 2.0 * K * 8 * V_ * V_ * O_
        rate = f_cpu / s_cpu
        out["cpu"] = {"value_s": flops / rate, "unit": "s (extrapolated, all blocks)",
                      "gflops": rate / 1e9, "threads": os.cpu_count(),
                      "sample": "(ab|ci) for 8 of the V values of a, plus (ai|bj) in "
                                f"{host_vvoo_s:.1f}s, numpy on the host cores"}
        out["speedup_vs_cpu"] = out["cpu"]["value_s"] / gpu_s
    return out


class CpuBaseline(threading.Thread):
    """Times the oracle on a seeded, size-stratified sample of the workload's
    tasks (n_per_quartile drawn from each flop quartile, smallest first) in a
    host thread beside the GPU work, which leaves the host cores idle while
    the device runs.  The oracle runs on all host cores but one (the one the
    GPU-driving thread uses).  A task starts only if it is predicted, at the
    rate measured so far, to end before the deadline (seconds from process
    start), and only while the timed steps have not ended (`stop`); at least
    one task runs.  Each finished task prints a line."""

    def __init__(self, tiling, s, per_task, deadline_s, n_per_quartile=2):
        super().__init__(daemon=True)
        self.tiling, self.s, self.per_task, self.deadline = tiling, s, per_task, deadline_s
        order = np.argsort(per_task, kind="stable")
        rng = np.random.default_rng(SEED)
        strata = np.array_split(order, 4)
        sample = np.concatenate([rng.choice(q, min(n_per_quartile, len(q)), replace=False)
                                 for q in strata])
        self.stratum = {int(k): i for i, q in enumerate(strata) for k in q}
        self.sample = sample[np.argsort(per_task[sample], kind="stable")]
        self.threads = max(1, (os.cpu_count() or 2) - 1)
        self.pick, self.parts, self.secs, self.error = [], [], [], None
        self.stop = threading.Event()

    def run(self):
        try:
            from oracle import oracle as O
            O.set_num_threads(self.threads)
            occ, vir = self.tiling.occ, self.tiling.vir
            t = O.Tiling(occ, occ, vir, vir)
            evl = np.concatenate([self.s["eo"], self.s["eo"], self.s["ev"], self.s["ev"]])
            rate = 5e9  # prior (flop/s) until the first task is timed
            for k in self.sample:
                if self.pick:
                    rate = float(self.per_task[self.pick].sum()) / sum(self.secs)
                    if self.stop.is_set() or elapsed() + float(self.per_task[k]) / rate > self.deadline:
                        break
                t0 = time.time()
                _, _, pt = O.run_spatial(t, evl, self.s, True, [int(k)])
                self.secs.append(time.time() - t0)
                self.pick.append(int(k))
                self.parts.append(pt[0])
                log(f"cpu baseline: task {int(k)} ({float(self.per_task[k]):.3e} flops) in "
                    f"{self.secs[-1]:.1f} s")
        except Exception as e:  # reported in the line, never silently
            self.error = repr(e)

    def result(self):
        spent, flops = float(sum(self.secs)), float(self.per_task[self.pick].sum())
        model = "unknown CPU"
        try:
            with open("/proc/cpuinfo") as f:
                model = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
        except (OSError, StopIteration):
            pass
        quart = [self.stratum[k] for k in self.pick]
        return dict(rate=flops / spent, spent=spent, pick=self.pick, flops=flops,
                    sampled=len(self.sample), task_flops=[float(self.per_task[k]) for k in self.pick],
                    task_s=self.secs, quartiles=quart, threads=self.threads,
                    partials=np.asarray(self.parts), model=model, nproc=os.cpu_count())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--budget-s", type=float, default=500.0,
                    help="wall-clock budget of the whole run from process start; warm-up and "
                         "timed steps are capped to fit it (at least one timed step runs)")
    ap.add_argument("--config", default="c3", choices=sorted(CONFIGS))
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--schedule", default="learned", choices=["driver", "analytic", "learned"])
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default ticket group shape instead of tuning it")
    ap.add_argument("--tune-sample", type=int, default=128,
                    help="tasks timed per candidate by the autotuner")
    ap.add_argument("--learn-sample", type=int, default=256,
                    help="tasks timed per precision to train the cost model")
    ap.add_argument("--no-graph", action="store_true", help="launch eagerly instead of the graph")
    ap.add_argument("--task-sample", type=int, default=0,
                    help="profiling aid: time only this many tasks (not the benchmark workload)")
    ap.add_argument("--compare-frac", type=float, default=0.02,
                    help="fraction of tasks in the seeded FP64 comparison sample")
    ap.add_argument("--fp64-full", action="store_true",
                    help="compare against a full FP64 run instead of a task sample")
    ap.add_argument("--no-compare", action="store_true")
    ap.add_argument("--no-schedule-ab", action="store_true",
                    help="skip timing the driver-order and analytic schedules on the sample")
    ap.add_argument("--sigma", type=float, default=0.05,
                    help="standard deviation of the synthetic integrals (E4 scales as sigma^4)")
    ap.add_argument("--t1-sigma", type=float, default=1e-3,
                    help="standard deviation of the synthetic T1 amplitudes")
    ap.add_argument("--no-physical", action="store_true",
                    help="skip the physical-scale check against the committed oracle fixture")
    ap.add_argument("--gather-tasks", type=int, default=16,
                    help="tasks of the device-side gather leg (0 skips it)")
    ap.add_argument("--gather-cpu-tasks", type=int, default=4,
                    help="of those, tasks also gathered by the oracle on one host core")
    ap.add_argument("--dropin-tasks", type=int, default=8,
                    help="tasks of the per-task drop-in leg: device gather plus "
                         "xt_fully_fused_ccsd_t_gpu (0 skips it)")
    ap.add_argument("--cholesky-leg", type=int, default=0, metavar="KFAC",
                    help="also time the V2 construction from KFAC*(O+V) Cholesky vectors on the "
                         "GPU (xt_upload_cholesky) against numpy on the host cores; 0 skips it")
    ap.add_argument("--expand-leg", action="store_true",
                    help="also time the dense spin-orbital T1/T2 expansion (xt_expand_amplitudes)")
    ap.add_argument("--cpu-deadline-s", type=float, default=0.0,
                    help="seconds from process start by which the CPU baseline stops starting "
                         "tasks (default: the wall-clock budget less 30 s)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(sharding the task list over GPUs is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    import exachem_amd as xt
    PREC = {"fp64": xt.PREC_FP64, "mixed": xt.PREC_MIXED}
    prec = PREC[args.precision]
    SCHED = {"driver": xt.SCHED_DRIVER, "analytic": xt.SCHED_ANALYTIC,
             "learned": xt.SCHED_LEARNED}
    sched = SCHED[args.schedule]
    graph = 0 if args.no_graph else 1
    tiling = make_tiling(xt, args.config)
    O_, V_ = tiling.O, tiling.V
    t0 = time.time()
    s = xt.synth_closed_shell(O_, V_, SEED, args.sigma, args.t1_sigma)
    ctx = xt.CCSDTriples(0)
    ctx.set_tiling(tiling)
    ctx.upload(s)
    upload_s = time.time() - t0
    total_flops, per_task, parts = ctx.flops()
    n_tasks = ctx.num_tasks()
    # the CPU baseline runs on the host cores beside all the GPU work below
    cpu = None
    if not args.no_cpu_baseline:
        cpu = CpuBaseline(tiling, s, per_task,
                          args.cpu_deadline_s if args.cpu_deadline_s > 0 else args.budget_s - 30.0)
        cpu.start()
    # operand conversion of the chosen precision (one-time, like the upload)
    t0 = time.time()
    ctx.prepare(prec)
    operand_s = time.time() - t0
    # one-time calibration (untimed, like the upload): tile-shape autotuning of
    # the measured precision and the learned cost model
    tune = {}
    t0 = time.time()
    if not args.no_autotune:
        tune_precs = [(args.precision, prec)]
        if args.fp64_full and args.precision != "fp64":
            tune_precs.append(("fp64", xt.PREC_FP64))
        for name, p in tune_precs:
            best, times, cands = ctx.autotune(precision=p, n_sample=args.tune_sample)
            tune[name] = {"shape": cands[best], "candidates": cands,
                          "sample_ms": [round(x, 2) for x in times]}
            log(f"autotune {name}: shape {cands[best]} ({min(times):.1f} ms vs default "
                f"{times[0]:.1f} ms on {args.tune_sample} tasks)")
    tune_s = time.time() - t0
    learn = None
    t0 = time.time()
    if args.schedule == "learned":
        model = xt.CostModel(0, seed=1)
        learn = ctx.learn_schedule(model, prec, n_sample=args.learn_sample)
        model.close()
        log(f"learned schedule: {learn}")
    learn_s = time.time() - t0
    which = None
    if args.task_sample:
        rng = np.random.default_rng(1)
        which = np.sort(rng.choice(n_tasks, min(args.task_sample, n_tasks), replace=False))
        total_flops = float(per_task[which].sum())
    log(f"{args.config} O={O_} V={V_} tasks={n_tasks} flops={total_flops:.4e} "
        f"upload={upload_s:.1f}s operands={operand_s:.1f}s occ={tiling.occ} vir={tiling.vir}")

    def step():
        if which is None:
            return ctx.run(precision=prec, schedule=sched, use_graph=graph, partials=False,
                           progress_ms=30000)
        return ctx.run_tasks(which, precision=prec, schedule=sched, use_graph=graph,
                             progress_ms=30000)

    line = {
        "metric": "time-to-E(T) + MFMA TFLOP/s at O=48,V=384; |dE(T)| vs FP64 oracle",
        "unit": "s", "n_gpus": 1, "higher_is_better": False, "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f16x3 products / f32 accum / f64 energy",
        "data": "synthetic (seeded generator, stated distributions)",
        "config": {"workload": f"closed-shell (T) energy, O={O_} V={V_}, "
                               f"{CONFIGS[args.config][2][0]} tiles, all tasks",
                   "O": O_, "V": V_, "occ_tiles": tiling.occ, "vir_tiles": tiling.vir,
                   "tasks": n_tasks if which is None else int(len(which)),
                   "task_sample": bool(which is not None), "precision": args.precision,
                   "schedule": args.schedule, "graph": bool(graph),
                   "sigma": args.sigma, "t1_sigma": args.t1_sigma,
                   "parallelism": "single GPU, persistent kernel over per-XCD work queues"},
    }

    # FP64 against the measured precision on a seeded task sample; its mixed
    # run under the chosen schedule also gives the estimate of one full step
    est_step = None
    if not args.no_compare and args.precision != "fp64":
        if args.fp64_full:
            cw = np.arange(n_tasks)
        else:
            rng = np.random.default_rng(SEED)
            cw = np.sort(rng.choice(n_tasks, max(1, int(args.compare_frac * n_tasks)),
                                    replace=False))
        log(f"comparison on {len(cw)} tasks")
        f = ctx.run_tasks(cw, precision=xt.PREC_FP64, schedule=sched, use_graph=graph,
                          progress_ms=30000)
        m = ctx.run_tasks(cw, precision=prec, schedule=sched, use_graph=graph, progress_ms=30000)
        d4, d5 = abs(m["e4"] - f["e4"]), abs(m["e5"] - f["e5"])
        cflops = float(per_task[cw].sum())
        est_step = m["kernel_ms"] / 1e3 * total_flops / cflops
        line["fp64_compare"] = {
            "tasks": int(len(cw)), "full": bool(args.fp64_full), "flops": cflops,
            "reference": "GPU FP64 kernel on the same tasks (FP64 matches the oracle to "
                         "1e-10 Eh + 2e-14 |E| per task, see oracle_check)",
            "fp64_kernel_ms": f["kernel_ms"], "kernel_ms": m["kernel_ms"],
            "speedup_vs_fp64": f["kernel_ms"] / m["kernel_ms"],
            "E4_fp64": f["e4"], "E5_fp64": f["e5"],
            "abs_dE4": d4, "abs_dE5": d5, "abs_dET": d5,
            "rel_dE4": d4 / abs(f["e4"]), "rel_dE5": d5 / abs(f["e5"]),
            "max_task_rel_dE4": float(np.max(np.abs(m["task_partials"][:, 0] -
                                                    f["task_partials"][:, 0]) /
                                             np.abs(f["task_partials"][:, 0]))),
        }
        log(f"fp64 {f['kernel_ms']:.1f} ms, {args.precision} {m['kernel_ms']:.1f} ms, "
            f"rel dE4 {d4 / abs(f['e4']):.2e}")
        if not args.no_schedule_ab:
            ab = {args.schedule: m["kernel_ms"]}
            for name in ("driver", "analytic", "learned"):
                if name not in ab and (name != "learned" or learn is not None):
                    r = ctx.run_tasks(cw, precision=prec, schedule=SCHED[name], use_graph=graph)
                    ab[name] = r["kernel_ms"]
                    if r["e4"] != m["e4"] or r["e5"] != m["e5"]:
                        raise AssertionError(f"schedule {name} changed the energies")
            line["schedule_ab"] = {
                "tasks": int(len(cw)), "precision": args.precision,
                "kernel_ms": ab, "energies_bit_identical": True,
                "learned_vs_analytic": ab.get("analytic", np.nan) / ab.get("learned", np.nan),
                "learned_vs_driver": ab.get("driver", np.nan) / ab.get("learned", np.nan)}
            log(f"schedule A/B: {ab}")

    # GPU against the oracle on the committed fixtures of this workload
    fx = {"c3": ("parity_c3_sample.json", "parity_c3_phys_sample.json"),
          "c4": ("parity_c4_sample.json", None)}.get(args.config, (None, None))
    precs = [("fp64", xt.PREC_FP64)] + ([(args.precision, prec)] if args.precision != "fp64"
                                         else [])
    case = golden_case(fx[0]) if fx[0] else None
    if (case and case["sigma"] == args.sigma and case["t1_sigma"] == args.t1_sigma
            and case["occ_tiles"] == tiling.occ and case["vir_tiles"] == tiling.vir):
        case["file"] = fx[0]
        line["oracle_check"] = fixture_check(xt, ctx, case, precs)
        log(f"oracle check: {line['oracle_check']}")

    if args.cholesky_leg > 0:
        line["cholesky_v2"] = cholesky_leg(xt, tiling, args.cholesky_leg,
                                           cpu=not args.no_cpu_baseline)
        log(f"cholesky v2: {line['cholesky_v2']}")

    if args.expand_leg:
        line["t_expansion"] = expand_leg(ctx, tiling)
        log(f"t expansion: {line['t_expansion']}")

    if args.gather_tasks > 0:
        line["device_gather"] = gather_leg(xt, ctx, tiling, s, args.gather_tasks,
                                           cpu_tasks=0 if args.no_cpu_baseline
                                           else args.gather_cpu_tasks)
        log(f"device gather: {line['device_gather']}")

    if args.dropin_tasks > 0:
        line["dropin_per_task"] = dropin_leg(xt, ctx, tiling, s, per_task, args.dropin_tasks)
        log(f"per-task drop-in: {line['dropin_per_task']}")

    del s

    # physical energy scale: the same tiling with integrals scaled down so
    # that |E(T)| is below 1 Eh, against its committed oracle fixture
    case = golden_case(fx[1]) if fx[1] and not args.no_physical else None
    if case and case["occ_tiles"] == tiling.occ and case["vir_tiles"] == tiling.vir:
        case["file"] = fx[1]
        t0 = time.time()
        pctx = xt.CCSDTriples(0)
        pctx.set_tiling(tiling)
        pctx.synth_upload(case["seed"], case["sigma"], case["t1_sigma"])
        if args.precision != "fp64":
            pctx.prepare(prec)
        line["oracle_check_physical"] = fixture_check(xt, pctx, case, precs)
        pctx.close()
        line["oracle_check_physical"]["seconds"] = time.time() - t0
        log(f"physical-scale check: {line['oracle_check_physical']}")

    # timed steps: as many of those requested as fit the budget (at least one)
    t_steps = elapsed()
    if est_step is None:
        est_step = max(1e-3, float(np.mean(
            [ctx.run_tasks(np.arange(min(n_tasks, 64)), precision=prec)["kernel_ms"]])) / 1e3 *
            total_flops / float(per_task[:min(n_tasks, 64)].sum()))
    fit = int(max(0.0, args.budget_s - t_steps) // (1.05 * est_step + 0.5))
    n_steps = min(args.steps, max(1, fit))
    n_warm = min(args.warmup, max(0, fit - n_steps))
    log(f"{t_steps:.0f}s elapsed; estimated step {est_step:.1f}s; budget {args.budget_s:.0f}s "
        f"fits {fit}: warm-up {n_warm}/{args.warmup}, steps {n_steps}/{args.steps}")
    for w in range(n_warm):
        step()
        log(f"warmup {w} done")
    times, kms, res = [], [], None
    for k in range(n_steps):
        t0 = time.perf_counter()
        res = step()
        times.append(time.perf_counter() - t0)
        kms.append(res["kernel_ms"])
        log(f"step {k}: {times[-1]:.3f}s kernel {res['kernel_ms']:.1f} ms "
            f"E4={res['e4']!r} E5={res['e5']!r}")
        if k + 1 < n_steps and elapsed() + 1.05 * times[-1] > args.budget_s:
            log("budget reached")
            break
    if cpu is not None:
        cpu.stop.set()  # no new task after the timed steps; the running one ends
        cpu.join()
        if cpu.error or not cpu.pick:
            raise RuntimeError(f"CPU baseline failed: {cpu.error}")
        cb = cpu.result()
        line["cpu_baseline"] = {
            "value": total_flops / cb["rate"], "unit": "s (extrapolated time-to-E(T))",
            "cores": cb["threads"], "kind": "port",
            "sample": f"{len(cb['pick'])} of {cb['sampled']} seeded size-stratified tasks "
                      f"(two per flop quartile, smallest first; quartiles {cb['quartiles']}, "
                      f"{min(cb['task_flops']):.2e}-{max(cb['task_flops']):.2e} flops each, "
                      f"{cb['flops']:.3e} in all) timed in {cb['spent']:.1f}s = "
                      f"{cb['rate'] / 1e9:.2f} GFLOP/s on {cb['threads']} OpenMP threads "
                      f"({cb['model']}, nproc {cb['nproc']}) in a host thread beside the GPU "
                      f"work, extrapolated by the analytic flop count",
            "tasks_timed": len(cb["pick"]),
            "task_gflops": [round(fl / sec_ / 1e9, 3) for fl, sec_ in
                            zip(cb["task_flops"], cb["task_s"])]}
        chk = {}
        for name, p in precs:
            g = ctx.run_tasks(cb["pick"], precision=p)["task_partials"]
            ref = cb["partials"]
            chk[name] = {"max_abs_dE4": float(np.max(np.abs(g[:, 0] - ref[:, 0]))),
                         "max_abs_dE5": float(np.max(np.abs(g[:, 1] - ref[:, 1]))),
                         "max_rel_dE4": float(np.max(np.abs(g[:, 0] - ref[:, 0]) /
                                                     np.abs(ref[:, 0])))}
            if p == xt.PREC_FP64:
                chk[name]["within_1e-10_plus_2e-14_rel"] = bool(
                    np.all(np.abs(g - ref) <= 1e-10 + 2e-14 * np.abs(ref)))
        line["cpu_sample_check"] = dict(tasks=cb["pick"], reference="oracle timed in this run",
                                        **chk)
        log(f"cpu baseline: {line['cpu_baseline']}")
    sec = float(np.mean(times))
    kernel_s = float(np.mean(kms)) / 1e3
    achieved = total_flops / kernel_s / 1e12
    peak = PEAK_TFLOPS[args.precision]
    line.update({
        "value": sec, "steps": len(times), "warmup": n_warm, "ms_per_step": sec * 1e3,
        "steps_requested": args.steps, "warmup_requested": args.warmup,
        "step_times_s": [round(x, 3) for x in times],
        "budget": {"budget_s": args.budget_s, "elapsed_before_steps_s": t_steps,
                   "estimated_step_s": est_step,
                   "warmup_note": "autotuner, cost-model and comparison launches of the same "
                                  "kernel and graph precede the timed steps"},
        "E4": res["e4"], "E5": res["e5"], "tflops": achieved,
        "setup_s": {"upload_and_layout": upload_s, "operand_split": operand_s,
                    "autotune": tune_s, "learn_schedule": learn_s},
        "autotune": tune or None,
        "learned_schedule": learn,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None,
                     "traffic_note": "HBM bytes per launch come from rocprofv3 --pmc passes "
                                     "outside this run (profiles/)",
                     "flops_per_launch": total_flops, "flop_count": "reference analytic count",
                     "kernel_ms": kernel_s * 1e3},
    })
    if args.precision == "mixed":
        # three FP16 MFMA products per d1/d2 multiply-add
        issued = (3 * (parts[1] + parts[2]) + parts[0]) if which is None else 3 * total_flops
        line["roofline"]["mfma_issued_tflops"] = issued / kernel_s / 1e12
    line["wall_s"] = elapsed()
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
