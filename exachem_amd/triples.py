"""Host-side front end of the fused (T) path.

``CCSDTriples`` mirrors the call sequence of ExaChem's
``CCSD_T_Fused_Driver<T>::execute`` (exachem/cc/ccsd_t/ccsd_t_fused_driver.hpp:
91-546): a tiled MO space with k_spin (ccsd_t.cpp:245-249), T1/T2 and V2
inputs, ``is_restricted``; it returns (E[4], E[5]) = (energy1, energy2) of
the reference's tuple, i.e. the [T] and (T) corrections.  All work happens in
the native library; errors surface as ``XtError`` carrying the library's
message (the reference would ``exit(100)`` instead, ccsd_t_common.hpp:35-41).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import MLP_IN, MLP_PARAMS, PREC_FP64, PREC_MIXED, SCHED_DRIVER, LearnStats, XtOpts, check, lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


class CostModel:
    """Task cost model (7 -> 64 -> 64 -> 1 FP32 perceptron) whose training and
    inference run as HIP kernels on `device`.  Parameters: flat float32 array
    W1[7][64] b1[64] W2[64][64] b2[64] w3[64] b3."""

    def __init__(self, device: int = 0, seed: int = 1):
        self._m = C.c_void_p()
        check(lib().xt_mlp_create(device, seed, C.byref(self._m)))

    def close(self):
        if self._m:
            lib().xt_mlp_destroy(self._m)
            self._m = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._m

    def get_params(self):
        p = np.zeros(MLP_PARAMS, dtype=np.float32)
        check(lib().xt_mlp_get_params(self._m, _fp(p)))
        return p

    def set_params(self, p):
        p = _f32(p, (MLP_PARAMS,))
        check(lib().xt_mlp_set_params(self._m, _fp(p)))

    def loss_grad(self, x, t):
        x = _f32(x)
        t = _f32(t)
        n = len(t)
        if x.shape != (n, MLP_IN):
            raise ValueError("x must be (n, 7)")
        loss = np.zeros(1, dtype=np.float32)
        g = np.zeros(MLP_PARAMS, dtype=np.float32)
        check(lib().xt_mlp_loss_grad(self._m, n, _fp(x), _fp(t), _fp(loss), _fp(g)))
        return float(loss[0]), g

    def train(self, x, t, steps, lr):
        x = _f32(x)
        t = _f32(t)
        n = len(t)
        if x.shape != (n, MLP_IN):
            raise ValueError("x must be (n, 7)")
        hist = np.zeros(max(steps, 1), dtype=np.float32)
        check(lib().xt_mlp_train(self._m, n, _fp(x), _fp(t), steps, lr, _fp(hist)))
        return hist[:steps]

    def predict(self, x):
        x = _f32(x)
        y = np.zeros(len(x), dtype=np.float32)
        check(lib().xt_mlp_predict(self._m, len(x), _fp(x), _fp(y)))
        return y


def uniform_tiles(n: int, ts: int):
    """Tiling of one spin part as setupMOIS does it (full tiles, remainder last;
    exachem/cholesky/cholesky_2e.cpp:226-277)."""
    return [ts] * (n // ts) + ([n % ts] if n % ts else [])


def uneven_tiles(n: int, lo: int, hi: int, seed: int):
    """Seeded uneven tiling: sizes drawn uniformly in [lo, hi]; the last tiles
    are adjusted so that every tile stays in [lo, hi] whenever n >= lo."""
    rng = np.random.default_rng(seed)
    out, rem = [], n
    while rem > 0:
        if rem <= hi:
            t = rem
        else:
            t = int(rng.integers(lo, hi + 1))
            if 0 < rem - t < lo:
                t = rem - lo
        out.append(t)
        rem -= t
    return out


@dataclass
class Tiling:
    O: int
    V: int
    occ: list = field(default_factory=list)
    vir: list = field(default_factory=list)

    @classmethod
    def uniform(cls, O, V, occ_ts, vir_ts):
        return cls(O, V, uniform_tiles(O, occ_ts), uniform_tiles(V, vir_ts))

    @classmethod
    def uneven(cls, O, V, occ_rng=(6, 20), vir_rng=(16, 48), seed=0):
        return cls(O, V, uneven_tiles(O, *occ_rng, seed), uneven_tiles(V, *vir_rng, seed + 1))

    def spin_orbital(self):
        """(tiles, k_spin) of |occ_a|occ_b|virt_a|virt_b| (beta mirrors alpha)."""
        tiles = self.occ + self.occ + self.vir + self.vir
        no, nv = len(self.occ), len(self.vir)
        spins = [1] * no + [2] * no + [1] * nv + [2] * nv
        return tiles, spins

    def so_energies(self, eo, ev):
        """k_evl_sorted of the reference (ccsd_t.cpp:209) in spin-orbital tile order."""
        return np.concatenate([eo, eo, ev, ev])


def synth_closed_shell(O, V, seed, sigma=0.05, t1_sigma=1e-3):
    """Seeded synthetic closed-shell inputs from the native generator (host arrays)."""
    out = dict(eo=np.zeros(O), ev=np.zeros(V), t1=np.zeros((V, O)), t2=np.zeros((V, V, O, O)),
               vvoo=np.zeros((V, V, O, O)), ovoo=np.zeros((O, V, O, O)),
               vvvo=np.zeros((V, V, V, O)))
    check(lib().xt_synth_closed_shell(O, V, seed, sigma, t1_sigma,
                                      *[_dp(out[k]) for k in
                                        ("eo", "ev", "t1", "t2", "vvoo", "ovoo", "vvvo")]))
    return out


def sum_task_energies(task_partials):
    """(E4, E5) from per-task contributions (n x 2, task-list order, factor
    included) in the fixed order of xt_run's device reduction: the partials
    of a run completed piecewise (xt_run_tasks over the remaining tasks) give
    the bit-identical totals of one full run (xt_sum_task_energies)."""
    p = np.ascontiguousarray(task_partials, dtype=np.float64).reshape(-1, 2)
    e = np.zeros(2)
    check(lib().xt_sum_task_energies(len(p), _dp(p), _dp(e)))
    return float(e[0]), float(e[1])


def default_opts(**kw):
    o = XtOpts()
    lib().xt_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class CCSDTriples:
    """One native context (one per host thread, one device)."""

    def __init__(self, device: int = 0):
        self._ctx = C.c_void_p()
        check(lib().xt_create(device, C.byref(self._ctx)))
        self.tiling = None

    def close(self):
        if self._ctx:
            lib().xt_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tiling(self, tiling: Tiling, restricted: bool = True):
        occ = np.ascontiguousarray(tiling.occ, dtype=np.int32)
        vir = np.ascontiguousarray(tiling.vir, dtype=np.int32)
        check(lib().xt_set_tiling(self._ctx, tiling.O, tiling.V, len(occ), _ip(occ), len(vir),
                                  _ip(vir), int(restricted)))
        self.tiling = tiling

    def num_tasks(self):
        n = C.c_long()
        check(lib().xt_num_tasks(self._ctx, C.byref(n)))
        return n.value

    def task_list(self):
        n = self.num_tasks()
        t = np.zeros(6 * n, dtype=np.int64)
        f = np.zeros(n)
        check(lib().xt_task_list(self._ctx, t.ctypes.data_as(C.POINTER(C.c_long)), _dp(f)))
        return t.reshape(n, 6), f

    def flops(self):
        n = self.num_tasks()
        per = np.zeros(max(n, 1))
        tot = C.c_double()
        parts = np.zeros(3)
        check(lib().xt_flops(self._ctx, _dp(per), C.byref(tot), _dp(parts)))
        return tot.value, per[:n], parts

    def memory_required(self, nchol=0):
        """Device bytes of the inputs of the current tiling: dict with the
        resident FP64 operands, the upload staging (closed-shell upload, or
        the Cholesky upload with nchol vectors) and the FP16 split copies."""
        b = np.zeros(3)
        check(lib().xt_memory_required(self._ctx, nchol, _dp(b)))
        return {"operands": b[0], "staging": b[1], "split": b[2]}

    def upload_peak(self):
        """Largest device bytes the last upload held at once (its own allocations)."""
        b = C.c_double()
        check(lib().xt_upload_peak_bytes(self._ctx, C.byref(b)))
        return b.value

    def upload(self, s):
        arrs = [np.ascontiguousarray(s[k], dtype=np.float64)
                for k in ("eo", "ev", "t1", "t2", "vvoo", "ovoo", "vvvo")]
        check(lib().xt_upload_closed_shell(self._ctx, *[_dp(a) for a in arrs]))

    def upload_cholesky(self, eo, ev, t1, t2, chol):
        """Inputs with the V2 blocks built on the GPU from Cholesky vectors
        chol [O+V][O+V][nchol] (cholVpr order, symmetric in the orbital pair),
        in place of V2TensorSetup::setupV2Tensors (cholesky/v2tensors.cpp:52-98)."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (eo, ev, t1, t2)]
        L = np.ascontiguousarray(chol, dtype=np.float64)
        nmo = len(arrs[0]) + len(arrs[1])
        assert L.ndim == 3 and L.shape[0] == nmo and L.shape[1] == nmo, L.shape
        check(lib().xt_upload_cholesky(self._ctx, *[_dp(a) for a in arrs], L.shape[2], _dp(L)))

    def synth_upload(self, seed, sigma=0.05, t1_sigma=1e-3):
        check(lib().xt_synth_upload(self._ctx, seed, sigma, t1_sigma))

    def prepare(self, precision):
        """Builds the operand form of `precision` ahead of the first run."""
        check(lib().xt_prepare(self._ctx, precision))

    def run(self, precision=PREC_FP64, schedule=SCHED_DRIVER, use_graph=0, record_task_time=0,
            workgroups=0, partials=True, progress_ms=0):
        n = self.num_tasks()
        o = default_opts(precision=precision, schedule=schedule, use_graph=use_graph,
                         record_task_time=record_task_time, workgroups=workgroups,
                         progress_ms=progress_ms)
        e = np.zeros(2)
        tp = np.zeros(2 * max(n, 1)) if partials else None
        tm = np.zeros(max(n, 1), dtype=np.float32) if record_task_time else None
        check(lib().xt_run(self._ctx, C.byref(o), _dp(e), _dp(tp) if tp is not None else None,
                           tm.ctypes.data_as(C.POINTER(C.c_float)) if tm is not None else None))
        ms = C.c_double()
        check(lib().xt_last_kernel_ms(self._ctx, C.byref(ms)))
        return dict(e4=float(e[0]), e5=float(e[1]),
                    task_partials=tp[:2 * n].reshape(n, 2) if tp is not None else None,
                    task_ms=tm[:n] if tm is not None else None, kernel_ms=ms.value)

    def run_tasks(self, which, precision=PREC_FP64, schedule=SCHED_DRIVER, use_graph=0,
                  record_task_time=0, workgroups=0, progress_ms=0):
        """Runs the listed tasks only; partials come back in listed order."""
        which = np.ascontiguousarray(which, dtype=np.int64)
        n = len(which)
        o = default_opts(precision=precision, schedule=schedule, use_graph=use_graph,
                         record_task_time=record_task_time, workgroups=workgroups,
                         progress_ms=progress_ms)
        e = np.zeros(2)
        tp = np.zeros(2 * n)
        tm = np.zeros(n, dtype=np.float32)
        check(lib().xt_run_tasks(self._ctx, C.byref(o), n,
                                 which.ctypes.data_as(C.POINTER(C.c_long)), _dp(e), _dp(tp),
                                 tm.ctypes.data_as(C.POINTER(C.c_float))))
        ms = C.c_double()
        check(lib().xt_last_kernel_ms(self._ctx, C.byref(ms)))
        return dict(e4=float(e[0]), e5=float(e[1]), task_partials=tp.reshape(n, 2),
                    task_ms=tm if record_task_time else None, kernel_ms=ms.value)

    def task_features(self, precision=PREC_FP64):
        """Cost-model inputs of every task: (n, 7) float32."""
        f = np.zeros((max(self.num_tasks(), 1), MLP_IN), dtype=np.float32)
        check(lib().xt_task_features(self._ctx, precision, _fp(f)))
        return f[:self.num_tasks()]

    def set_task_cost(self, cost):
        """Per-task costs for SCHED_LEARNED."""
        cost = np.ascontiguousarray(cost, dtype=np.float64)
        check(lib().xt_set_task_cost(self._ctx, _dp(cost)))

    def learn_schedule(self, model: CostModel, precision=PREC_MIXED, n_sample=256,
                       precisions=(PREC_FP64, PREC_MIXED), steps=1500, lr=3e-3, workgroups=0):
        """Times n_sample tasks per precision, trains `model` on the GPU and
        installs its predicted costs for SCHED_LEARNED.  Returns the stats."""
        o = default_opts(precision=precision, workgroups=workgroups)
        mask = sum(1 << (0 if p == PREC_FP64 else 1) for p in precisions)
        st = LearnStats()
        check(lib().xt_learn_schedule(self._ctx, model.handle, C.byref(o), n_sample, mask, steps,
                                      lr, C.byref(st)))
        return {k: getattr(st, k) for k, _ in LearnStats._fields_}

    def set_group_shape(self, precision, glog):
        """Ticket group shape (log2 extents along h1,h2,h3,p4,p5,p6, sum 6)."""
        g = np.ascontiguousarray(glog, dtype=np.int32)
        check(lib().xt_set_group_shape(self._ctx, precision, _ip(g)))

    def group_shape(self, precision):
        g = np.zeros(6, dtype=np.int32)
        check(lib().xt_get_group_shape(self._ctx, precision, _ip(g)))
        return g.tolist()

    def autotune(self, precision=PREC_MIXED, n_sample=256, schedule=SCHED_DRIVER, workgroups=0):
        """Times every candidate group shape on a task sample and installs the
        fastest; returns (best index, times in ms, candidate shapes)."""
        o = default_opts(precision=precision, schedule=schedule, workgroups=workgroups)
        best = np.zeros(1, dtype=np.int32)
        times = np.zeros(6)
        shapes = np.zeros(36, dtype=np.int32)
        check(lib().xt_autotune(self._ctx, C.byref(o), n_sample, _ip(best), _dp(times),
                                _ip(shapes)))
        return int(best[0]), times.tolist(), shapes.reshape(6, 6).tolist()

    def packed_sizes(self):
        """Packed buffer sizes of the per-task path (doubles): s1_t1, s1_v2,
        d1_t2, d1_v2, d2_t2, d2_v2, and the slot counts max_d1, max_d2
        (ccsd_t_fused_driver.hpp:176-195)."""
        out = np.zeros(8, dtype=np.int64)
        check(lib().xt_packed_sizes(self._ctx, out.ctypes.data_as(C.POINTER(C.c_long))))
        names = ("s1_t1", "s1_v2", "d1_t2", "d1_v2", "d2_t2", "d2_v2", "max_d1", "max_d2")
        return {n: int(v) for n, v in zip(names, out)}

    def queue_layout(self, which=None, precision=PREC_FP64, schedule=SCHED_DRIVER):
        """The work queues a run over `which` (all tasks if None) would launch
        with: dict of qbeg[9], qtask, qtick, qtotal[8], nbox[ntasks][6] and
        the group shape glog[6].  Host only."""
        nt = self.num_tasks()
        w = (np.zeros(0, dtype=np.int64) if which is None
             else np.ascontiguousarray(which, dtype=np.int64))
        ne = nt if which is None else len(w)
        o = default_opts(precision=precision, schedule=schedule)
        qbeg = np.zeros(9, dtype=np.int32)
        qtask = np.zeros(max(ne, 1), dtype=np.int32)
        qtick = np.zeros(max(ne, 1), dtype=np.int64)
        qtotal = np.zeros(8, dtype=np.int64)
        nbox = np.zeros(6 * max(nt, 1), dtype=np.int32)
        glog = np.zeros(6, dtype=np.int32)
        ll = C.POINTER(C.c_longlong)
        check(lib().xt_queue_layout(self._ctx, C.byref(o), len(w),
                                    w.ctypes.data_as(C.POINTER(C.c_long)), _ip(qbeg), _ip(qtask),
                                    qtick.ctypes.data_as(ll), qtotal.ctypes.data_as(ll),
                                    _ip(nbox), _ip(glog)))
        return dict(qbeg=qbeg, qtask=qtask[:ne], qtick=qtick[:ne], qtotal=qtotal,
                    nbox=nbox[:6 * nt].reshape(nt, 6), glog=glog)

    def gather_task(self, tiles, bufs, stream=None):
        """Device-side gather of one task's packed buffers, in place of
        ccsd_t_data_{s1,d1,d2}_new (ccsd_t_all_fused.hpp:149-178).

        tiles: (t_h1b, t_h2b, t_h3b, t_p4b, t_p5b, t_p6b) spin-orbital tile
        indices; bufs: device pointers (int) keyed d1_t2, d1_v2, d2_t2, d2_v2,
        s1_t1, s1_v2, sized by packed_sizes(); stream: hipStream_t (int) or
        None for the null stream.  Returns (meta, counts): meta holds the
        host int32 arrays s1_size, s1_exec, d1_size, d1_exec, d2_size,
        d2_exec; counts the enabled s1, d1, d2 slots."""
        t = np.ascontiguousarray(tiles, dtype=np.int64)
        noab, nvab = 2 * len(self.tiling.occ), 2 * len(self.tiling.vir)
        meta = {"s1_size": np.zeros(6, np.int32), "s1_exec": np.zeros(9, np.int32),
                "d1_size": np.zeros(7 * noab, np.int32), "d1_exec": np.zeros(9 * noab, np.int32),
                "d2_size": np.zeros(7 * nvab, np.int32), "d2_exec": np.zeros(9 * nvab, np.int32)}
        counts = np.zeros(3, np.int32)
        names = ("d1_t2", "d1_v2", "d2_t2", "d2_v2", "s1_t1", "s1_v2")
        check(lib().xt_gather_task(
            self._ctx, t.ctypes.data_as(C.POINTER(C.c_long)), stream,
            *[C.c_void_p(int(bufs[n])) for n in names],
            *[_ip(meta[k]) for k in ("s1_size", "s1_exec", "d1_size", "d1_exec", "d2_size",
                                     "d2_exec")], _ip(counts)))
        return meta, [int(x) for x in counts]

    def expand_amplitudes(self, t1_ptr, t2_ptr, stream=None):
        """Dense spin-orbital T1 [2V][2O] and T2 [2V][2V][2O][2O] into device
        buffers (int pointers or None), the expansion of
        cd_ccsd_cs_ann.cpp:932-976."""
        check(lib().xt_expand_amplitudes(self._ctx, stream,
                                         C.c_void_p(int(t1_ptr)) if t1_ptr else None,
                                         C.c_void_p(int(t2_ptr)) if t2_ptr else None))

    def gather_last_elems(self):
        """Doubles written by the last gather_task."""
        n = C.c_long()
        check(lib().xt_gather_last_elems(self._ctx, C.byref(n)))
        return n.value

    def task_codes(self, task):
        """(d1code, d2code, s1code) bit fields of one task."""
        codes = (C.c_uint64 * 3)()
        check(lib().xt_task_codes(self._ctx, task, codes))
        return tuple(int(x) for x in codes)



def fully_fused_ccsd_t_gpu(stream, num_blocks, dims, dev_bufs, meta, noab, nvab, max_dims, evl,
                           partial, done_copy=None, noalloc=False):
    """Per-task drop-in for ``fully_fused_ccsd_t_gpu<double>``
    (ccsd_t_all_fused_nontcCuda_Hip_Sycl.cpp:2800-2817).

    dims: (h1,h2,h3,p4,p5,p6); dev_bufs: device pointers of
    (d1_t2, d1_v2, d2_t2, d2_v2, s1_t1, s1_v2); meta: dict of int32 host arrays
    d1_size, d1_exec, d2_size, d2_exec, s1_size, s1_exec; max_dims: dict with
    d1_t2, d1_v2, d2_t2, d2_v2, s1_t1, s1_v2 (doubles per slot); evl: 6 device
    pointers; partial: device pointer of 2*num_blocks doubles.  noalloc: the
    allocation-free entry point (xt_fully_fused_ccsd_t_gpu_noalloc)."""
    m = {k: np.ascontiguousarray(meta[k], dtype=np.int32) for k in
         ("d1_size", "d1_exec", "d2_size", "d2_exec", "s1_size", "s1_exec")}
    fn = lib().xt_fully_fused_ccsd_t_gpu_noalloc if noalloc else lib().xt_fully_fused_ccsd_t_gpu
    return check(fn(
        stream, num_blocks, *dims, *dev_bufs,
        _ip(m["d1_size"]), _ip(m["d1_exec"]), _ip(m["d2_size"]), _ip(m["d2_exec"]),
        _ip(m["s1_size"]), _ip(m["s1_exec"]),
        noab, max_dims["d1_t2"], max_dims["d1_v2"], nvab, max_dims["d2_t2"], max_dims["d2_v2"],
        max_dims["s1_t1"], max_dims["s1_v2"], *evl, partial, done_copy))
