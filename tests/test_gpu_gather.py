"""Device-side gather of the per-task packed buffers (xt_gather_task) against
the oracle's restatement of the reference's host gathers
ccsd_t_data_{s1,d1,d2}_new: buffers bit-identical, launcher metadata equal,
and the fused per-task kernel fed straight from the gathered device buffers
reproducing each task's oracle energy.

The gather only copies values, forms x - y for antisymmetrised elements and
multiplies by +-1, so the bar is bit equality; the buffers are zeroed first
on both sides so that the unwritten tail of every slot compares too."""
import json
import os

import numpy as np
import pytest

from tests.test_gpu_parity import close, load_case, small_cases

HERE = os.path.dirname(os.path.abspath(__file__))
NAMES = ("d1_t2", "d1_v2", "d2_t2", "d2_v2", "s1_t1", "s1_v2")


def _tasks_for(case, oracle):
    occ, vir = case["occ_tiles"], case["vir_tiles"]
    t = oracle.Tiling(occ, occ, vir, vir)
    tasks, fac = oracle.task_list(t, case["restricted"])
    return t, tasks, fac


def test_packed_sizes_match_driver(xt, oracle):
    """Host-only: the packed buffer sizes equal the driver's
    (ccsd_t_fused_driver.hpp:176-195 as the oracle restates it)."""
    for case in small_cases():
        t, _, _ = _tasks_for(case, oracle)
        ctx = xt.CCSDTriples(-1)
        ctx.set_tiling(xt.Tiling(case["O"], case["V"], case["occ_tiles"], case["vir_tiles"]),
                       case["restricted"])
        assert ctx.packed_sizes() == oracle.buffer_sizes(t), case["name"]
        ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["uneven_small", "uneven_small_unrestricted", "config0",
                                  "wide_tiles"])
def test_device_gather_matches_host_gather(xt, oracle, name):
    from exachem_amd.hipmem import DevBuf, Stream
    case = next(c for c in small_cases() if c["name"] == name)
    ctx, s = load_case(xt, case)
    t, tasks, fac = _tasks_for(case, oracle)
    so = oracle.expand_closed_shell(case["O"], case["V"], s["t1"], s["t2"], s["vvoo"], s["ovoo"],
                                    s["vvvo"])
    sz = ctx.packed_sizes()
    dev = {n: DevBuf(8 * sz[n]) for n in NAMES}
    st = Stream()
    evl = np.concatenate([s["eo"], s["eo"], s["ev"], s["ev"]])
    offs = np.concatenate([[0], np.cumsum(t.tiles)])
    ref = np.asarray(case["task_partials"])
    ks = range(len(tasks)) if len(tasks) <= 40 else list(range(0, len(tasks), 3))
    for k in ks:
        print(f"[gpu] gather {name} task {k} tiles {list(tasks[k])}", flush=True)
        bufs, meta, counts, _ = oracle.task_packed(t, so, tasks[k], case["restricted"])
        for b in dev.values():
            b.zero()
        m, c = ctx.gather_task(tasks[k], {n: dev[n].ptr.value for n in NAMES}, st.h)
        st.sync()
        assert tuple(c) == counts, (k, c, counts)
        for key in meta:
            assert np.array_equal(m[key], meta[key]), (k, key, m[key], meta[key])
        for n in NAMES:
            got = dev[n].to_array(sz[n])
            assert np.array_equal(got.view(np.uint64), bufs[n].view(np.uint64)), (k, n)
        # the fused kernel reads the gathered device buffers directly
        dims = [int(t.tiles[b]) for b in tasks[k]]
        ev = [DevBuf.from_array(evl[offs[b]:offs[b] + t.tiles[b]]) for b in tasks[k]]
        nb = int(np.prod([(d + 3) // 4 for d in dims]))
        part = DevBuf(16 * nb)
        max_dims = {n: sz[n] // (9 if n.startswith("s1") else
                                 (sz["max_d1"] if n.startswith("d1") else sz["max_d2"]))
                    for n in NAMES}
        xt.fully_fused_ccsd_t_gpu(st.h, nb, dims, [dev[n].ptr for n in NAMES], m, t.noab, t.nvab,
                                  max_dims, [e.ptr for e in ev], part.ptr)
        st.sync()
        p = part.to_array(2 * nb)
        e4, e5 = fac[k] * np.sum(p[:nb]), fac[k] * np.sum(p[nb:])
        for b in ev + [part]:
            b.free()
        assert close(e4, ref[k, 0]) and close(e5, ref[k, 1]), (k, e4, ref[k, 0])
    for b in dev.values():
        b.free()
    st.destroy()
    ctx.close()


def test_spatial_host_gather_matches_dense(xt, oracle):
    """CPU: the oracle's host gather from spatial inputs (the CPU side that
    bench.py times) writes the same buffers and metadata as its dense
    spin-orbital form, which the GPU tests above pin against the device."""
    case = next(c for c in small_cases() if c["name"] == "uneven_small")
    s = xt.synth_closed_shell(case["O"], case["V"], case["seed"], case["sigma"], case["t1_sigma"])
    t, tasks, _ = _tasks_for(case, oracle)
    so = oracle.expand_closed_shell(case["O"], case["V"], s["t1"], s["t2"], s["vvoo"], s["ovoo"],
                                    s["vvvo"])
    g = oracle.SpatialGather(t, s, case["restricted"])
    for k in range(0, len(tasks), 5):
        for b in g.bufs.values():
            b[:] = 0.0
        n = g(tasks[k])
        bufs, meta, counts, _ = oracle.task_packed(t, so, tasks[k], case["restricted"])
        assert tuple(int(c) for c in g.counts) == counts
        assert n > 0
        for key in bufs:
            assert np.array_equal(g.bufs[key].view(np.uint64), bufs[key].view(np.uint64)), (k, key)
        m = np.concatenate([meta[key] for key in ("s1_size", "s1_exec", "d1_size", "d1_exec",
                                                  "d2_size", "d2_exec")])
        assert np.array_equal(g.meta, m), k


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["uneven_small", "config0"])
def test_amplitude_expansion_matches_oracle(xt, oracle, name):
    """xt_expand_amplitudes (cd_ccsd_cs_ann.cpp:932-976) writes the dense
    spin-orbital T1/T2 the oracle's expansion writes, bit for bit."""
    from exachem_amd.hipmem import DevBuf, Stream
    case = next(c for c in small_cases() if c["name"] == name)
    ctx, s = load_case(xt, case)
    so = oracle.expand_closed_shell(case["O"], case["V"], s["t1"], s["t2"], s["vvoo"], s["ovoo"],
                                    s["vvvo"])
    n1, n2 = so[0].size, so[1].size
    b1, b2 = DevBuf(8 * n1), DevBuf(8 * n2)
    st = Stream()
    print(f"[gpu] expand {name}: t1 {so[0].shape} t2 {so[1].shape}", flush=True)
    ctx.expand_amplitudes(b1.ptr.value, b2.ptr.value, st.h)
    st.sync()
    assert np.array_equal(b1.to_array(n1).view(np.uint64), so[0].ravel().view(np.uint64))
    assert np.array_equal(b2.to_array(n2).view(np.uint64), so[1].ravel().view(np.uint64))
    for b in (b1, b2):
        b.free()
    st.destroy()
    ctx.close()
