"""The allocation-free per-task launcher (xt_fully_fused_ccsd_t_gpu_noalloc)
against the regular one (xt_fully_fused_ccsd_t_gpu), through the C ABI: two
calls in a row on one stream, the second on a smaller task than the first
(the device scratch keeps the first call's metadata, which must not leak
into the second), each compared bit for bit with the regular entry point on
the same packed buffers and, after the host-side reduction of
hostEnergyReduce, with the oracle's task energy (1e-10 Eh + 2e-14 |E|).
Also: the caller's host arrays are reusable once the call returns, and the
same limits are refused."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
pytestmark = pytest.mark.gpu

NAMES = ("d1_t2", "d1_v2", "d2_t2", "d2_v2", "s1_t1", "s1_v2")


def _case(name):
    cases = json.load(open(os.path.join(HERE, "golden", "parity_small.json")))
    return next(c for c in cases if c["name"] == name)


def _task_inputs(xt, oracle, case, k):
    from exachem_amd.hipmem import DevBuf
    s = xt.synth_closed_shell(case["O"], case["V"], case["seed"], case["sigma"], case["t1_sigma"])
    occ, vir = case["occ_tiles"], case["vir_tiles"]
    t = oracle.Tiling(occ, occ, vir, vir)
    so = oracle.expand_closed_shell(case["O"], case["V"], s["t1"], s["t2"], s["vvoo"], s["ovoo"],
                                    s["vvvo"])
    tasks, fac = oracle.task_list(t, case["restricted"])
    bufs, meta, _, sz = oracle.task_packed(t, so, tasks[k], case["restricted"])
    evl = np.concatenate([s["eo"], s["eo"], s["ev"], s["ev"]])
    offs = np.concatenate([[0], np.cumsum(t.tiles)])
    dims = [int(t.tiles[b]) for b in tasks[k]]
    nb = int(np.prod([(d + 3) // 4 for d in dims]))
    dev = {n: DevBuf.from_array(bufs[n]) for n in NAMES}
    ev = [DevBuf.from_array(evl[offs[b]:offs[b] + t.tiles[b]]) for b in tasks[k]]
    max_dims = {n: sz[n] // (9 if n.startswith("s1") else
                             (sz["max_d1"] if n.startswith("d1") else sz["max_d2"])) for n in NAMES}
    return dict(dims=dims, nb=nb, dev=dev, ev=ev, meta=meta, max_dims=max_dims, noab=t.noab,
                nvab=t.nvab, factor=fac[k], ref=np.asarray(case["task_partials"])[k])


def _launch(xt, st, x, noalloc, part):
    meta = {k: np.array(v, dtype=np.int32, copy=True) for k, v in x["meta"].items()}
    xt.fully_fused_ccsd_t_gpu(st.h, x["nb"], x["dims"], [x["dev"][n].ptr for n in NAMES], meta,
                              x["noab"], x["nvab"], x["max_dims"], [e.ptr for e in x["ev"]],
                              part.ptr, noalloc=noalloc)
    for v in meta.values():  # the caller may overwrite its arrays once the call returns
        v[:] = -7


def test_noalloc_launcher_matches_regular_launcher(xt, oracle):
    from exachem_amd.hipmem import DevBuf, Stream
    st = Stream()
    # first call on a large task (one 50-wide virtual tile: 17,576 blocks,
    # 143 metadata ints), second on a smaller one (512 blocks, 79 ints)
    wide = _case("wide_tiles")
    occ, vir = wide["occ_tiles"], wide["vir_tiles"]
    t = oracle.Tiling(occ, occ, vir, vir)
    tasks, _ = oracle.task_list(t, wide["restricted"])
    big = max(range(len(tasks)), key=lambda k: np.prod([(t.tiles[b] + 3) // 4 for b in tasks[k]]))
    first = _task_inputs(xt, oracle, wide, big)
    small_case = _case("one_tile")
    second = _task_inputs(xt, oracle, small_case, 0)
    n_meta = [16 * x["noab"] + 16 * x["nvab"] + 15 for x in (first, second)]
    assert second["nb"] < first["nb"] and n_meta[1] < n_meta[0], (first["nb"], second["nb"], n_meta)
    outs = []
    for i, x in enumerate((first, second)):
        print(f"[gpu] no-alloc call {i}: dims {x['dims']} blocks {x['nb']}", flush=True)
        pa, pb = DevBuf(16 * x["nb"]), DevBuf(16 * x["nb"])
        _launch(xt, st, x, False, pa)
        _launch(xt, st, x, True, pb)
        st.sync()
        a, b = pa.to_array(2 * x["nb"]), pb.to_array(2 * x["nb"])
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), i
        e4 = x["factor"] * np.sum(b[:x["nb"]])
        e5 = x["factor"] * np.sum(b[x["nb"]:])
        tol = 1e-10 + 2e-14 * np.abs(x["ref"])
        assert abs(e4 - x["ref"][0]) <= tol[0] and abs(e5 - x["ref"][1]) <= tol[1], (i, e4, x["ref"])
        outs.append(b)
        for buf in [pa, pb] + list(x["dev"].values()) + x["ev"]:
            buf.free()
    # limits are refused as by the regular launcher
    x = _task_inputs(xt, oracle, small_case, 1)
    part = DevBuf(16 * x["nb"])
    for noalloc in (False, True):
        bad = dict(x, noab=51)
        with pytest.raises(xt.XtError) as e:
            _launch(xt, st, bad, noalloc, part)
        assert e.value.code == xt.XT_EINVAL
    part.free()
    st.destroy()
