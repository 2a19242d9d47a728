"""CPU tests of the product's host side: the C ABI loads and exports what the
header declares, the planner reproduces the reference driver's task list,
factors, slot enables and flop counts (checked against the oracle), the
generator keeps its promised symmetries, and errors come back as codes."""
import ctypes as C
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_library_exports_every_header_symbol(xt):
    syms = xt.header_symbols()
    assert len(syms) >= 15
    L = xt.lib()
    for s in syms:
        assert hasattr(L, s), s


def host_ctx(xt, O_, V_, occ, vir, restricted=True):
    ctx = xt.CCSDTriples(-1)
    ctx.set_tiling(xt.Tiling(O_, V_, list(occ), list(vir)), restricted)
    return ctx


CASES = [
    (9, 36, [9], [36]),                         # seh2 tiling
    (21, 14, [21], [14]),                       # butanol2 tiling
    (29, 103, [29], [40, 40, 23]),              # uracil tiling
    (8, 48, [8], [16, 16, 16]),                 # config 0
    (7, 22, [3, 4], [5, 9, 8]),                 # uneven
    (13, 40, [6, 7], [16, 13, 11]),             # uneven, config-3-like ranges
]


@pytest.mark.parametrize("O_,V_,occ,vir", CASES)
@pytest.mark.parametrize("restricted", [True, False])
def test_task_list_factors_flops_match_oracle(xt, oracle, O_, V_, occ, vir, restricted):
    ctx = host_ctx(xt, O_, V_, occ, vir, restricted)
    tasks, fac = ctx.task_list()
    t = oracle.Tiling(occ, occ, vir, vir)
    otasks, ofac = oracle.task_list(t, restricted)
    assert np.array_equal(tasks, otasks)
    assert np.array_equal(fac, ofac)
    tot, per, parts = ctx.flops()
    ototal, oparts, oper, _ = oracle.flops(t, restricted, per_task=True)
    assert tot == ototal and np.array_equal(per, oper)
    assert tuple(int(p) for p in parts) == oparts


def test_uracil_known_answer_through_product(xt):
    ctx = host_ctx(xt, 29, 103, [29], [40, 40, 23])
    assert ctx.flops()[0] == 62789556886348


def _decode(codes):
    d1 = [(codes[0] >> (6 * g)) & 63 for g in range(9)]
    d2 = [(codes[1] >> (6 * g)) & 63 for g in range(9)]
    s1 = [(codes[2] >> (4 * s)) & 15 for s in range(9)]
    return d1, d2, s1


@pytest.mark.parametrize("O_,V_,occ,vir", CASES[3:])
@pytest.mark.parametrize("restricted", [True, False])
def test_slot_enables_match_reference_exec_flags(xt, oracle, O_, V_, occ, vir, restricted):
    """Bipartition g = 3*hs + ps is fed by reference d1 slot 3*((ps+1)%3)+hs
    and d2 slot 3*ps+(hs+1)%3; s1 slot numbering is shared (3*pa+hi).  The
    product's enable bit must equal 'some h7 (p7) tile has exec >= 0' and the
    contraction must span every tile of one spin."""
    ctx = host_ctx(xt, O_, V_, occ, vir, restricted)
    t = oracle.Tiling(occ, occ, vir, vir)
    tasks, _ = ctx.task_list()
    spins = np.asarray(t.spins)
    for k in range(len(tasks)):
        d1, d2, s1 = _decode(ctx.task_codes(k))
        os1, od1, od2 = oracle.task_exec(t, tasks[k], restricted)
        for g in range(9):
            hs, ps = divmod(g, 3)
            r1 = 3 * ((ps + 1) % 3) + hs
            r2 = 3 * ps + (hs + 1) % 3
            on1 = od1[:, r1] >= 0
            on2 = od2[:, r2] >= 0
            assert bool(d1[g] & 1) == bool(on1.any()), (k, g)
            assert bool(d2[g] & 1) == bool(on2.any()), (k, g)
            if on1.any():  # all occupied tiles of exactly one spin
                sp = set(spins[:t.noab][on1])
                assert len(sp) == 1 and on1.sum() == np.sum(spins[:t.noab] == sp.pop())
            if on2.any():
                sp = set(spins[t.noab:][on2])
                assert len(sp) == 1 and on2.sum() == np.sum(spins[t.noab:] == sp.pop())
        for sl in range(9):
            assert bool(s1[sl] & 1) == bool(os1[sl] >= 0), (k, sl)


def test_generator_symmetry_and_determinism(xt):
    a = xt.synth_closed_shell(5, 11, 42)
    b = xt.synth_closed_shell(5, 11, 42)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(a["vvoo"], a["vvoo"].transpose(1, 0, 3, 2))
    assert np.array_equal(a["ovoo"], a["ovoo"].transpose(0, 1, 3, 2))
    assert np.array_equal(a["vvvo"], a["vvvo"].transpose(1, 0, 2, 3))
    assert np.array_equal(a["t2"], a["t2"].transpose(1, 0, 3, 2))
    D = (a["eo"][None, None, :, None] + a["eo"][None, None, None, :]) - (
        a["ev"][:, None, None, None] + a["ev"][None, :, None, None])
    assert np.allclose(a["t2"] * D, a["vvoo"], rtol=1e-15, atol=0)
    assert np.all(np.diff(a["eo"]) >= 0) and np.all(np.diff(a["ev"]) >= 0)
    assert a["eo"].min() >= -2.0 and a["eo"].max() <= -0.3
    assert a["ev"].min() >= 0.1 and a["ev"].max() <= 3.0
    c = xt.synth_closed_shell(5, 11, 43)
    assert not np.array_equal(a["vvvo"], c["vvvo"])
    big = xt.synth_closed_shell(6, 40, 1)
    assert abs(big["vvvo"].std() - 0.05) < 0.002 and abs(big["vvvo"].mean()) < 0.002


def test_generator_matches_golden_checksums(xt):
    cases = json.load(open(os.path.join(HERE, "golden", "parity_small.json")))
    for case in cases:
        s = xt.synth_closed_shell(case["O"], case["V"], case["seed"], case["sigma"],
                                  case["t1_sigma"])
        for k, (a, b) in case["checksums"].items():
            assert float(np.sum(s[k])) == a and float(np.sum(s[k] * s[k])) == b, (case["name"], k)


def test_error_codes_not_exits(xt):
    L = xt.lib()
    ctx = xt.CCSDTriples(-1)
    occ = np.array([3, 0], dtype=np.int32)
    vir = np.array([5], dtype=np.int32)
    rc = L.xt_set_tiling(ctx._ctx, 3, 5, 2, occ.ctypes.data_as(C.POINTER(C.c_int)), 1,
                         vir.ctypes.data_as(C.POINTER(C.c_int)), 1)
    assert rc == xt.XT_EINVAL and b"empty" in L.xt_last_error()
    occ = np.array([3], dtype=np.int32)
    rc = L.xt_set_tiling(ctx._ctx, 4, 5, 1, occ.ctypes.data_as(C.POINTER(C.c_int)), 1,
                         vir.ctypes.data_as(C.POINTER(C.c_int)), 1)
    assert rc == xt.XT_EINVAL
    with pytest.raises(xt.XtError) as ei:
        ctx.num_tasks()
    assert ei.value.code == xt.XT_ESTATE
    ctx.set_tiling(xt.Tiling(3, 5, [3], [5]))
    with pytest.raises(xt.XtError) as ei:
        ctx.run()
    assert ei.value.code == xt.XT_ENODEV
    # per-task drop-in: limits of the reference are errors, not overflows
    z = np.zeros(7 * 60, dtype=np.int32)
    zp = z.ctypes.data_as(C.POINTER(C.c_int))
    one = C.c_void_p(8)
    rc = L.xt_fully_fused_ccsd_t_gpu(None, 1, 4, 4, 4, 4, 4, 4, *[one] * 6, *[zp] * 6,
                                     51, 1, 1, 1, 1, 1, 1, 1, *[one] * 6, one, None)
    assert rc == xt.XT_EINVAL and b"noab" in L.xt_last_error()
    rc = L.xt_fully_fused_ccsd_t_gpu(None, 2, 4, 4, 4, 4, 4, 4, *[one] * 6, *[zp] * 6,
                                     2, 1, 1, 2, 1, 1, 1, 1, *[one] * 6, one, None)
    assert rc == xt.XT_EINVAL and b"num_blocks" in L.xt_last_error()
    rc = L.xt_fully_fused_ccsd_t_gpu(None, 1, 4, 4, 4, 4, 4, 4, *[None] * 6, *[zp] * 6,
                                     2, 1, 1, 2, 1, 1, 1, 1, *[one] * 6, one, None)
    assert rc == xt.XT_EINVAL
    # an exec slot the kernel's 30-bit tile-list field cannot hold is refused
    # before any device work
    d1s, d2s = np.zeros(7, np.int32), np.zeros(7, np.int32)
    d1s[3], d2s[6] = 4, 4
    d1e, d2e, s1e = np.full(9, -1, np.int32), np.full(9, -1, np.int32), np.full(9, -1, np.int32)
    d1e[0] = 1 << 30
    s1s = np.full(6, 4, np.int32)
    arrs = [a.ctypes.data_as(C.POINTER(C.c_int)) for a in (d1s, d1e, d2s, d2e, s1s, s1e)]
    rc = L.xt_fully_fused_ccsd_t_gpu(None, 1, 4, 4, 4, 4, 4, 4, *[one] * 6, *arrs,
                                     1, 1, 1, 1, 1, 1, 1, 1, *[one] * 6, one, None)
    assert rc == xt.XT_EINVAL and b"exec slot" in L.xt_last_error()


def test_group_shape_validation(xt):
    """Ticket group shapes are log2 extents summing to 6; others are rejected
    with an error code, not a crash (host-only context)."""
    ctx = host_ctx(xt, 7, 22, [3, 4], [5, 9, 8])
    assert ctx.group_shape(xt.PREC_MIXED) == [1, 1, 1, 1, 1, 1]
    ctx.set_group_shape(xt.PREC_MIXED, [0, 0, 0, 2, 2, 2])
    assert ctx.group_shape(xt.PREC_MIXED) == [0, 0, 0, 2, 2, 2]
    assert ctx.group_shape(xt.PREC_FP64) == [1, 1, 1, 1, 1, 1]
    for bad in ([1, 1, 1, 1, 1, 2], [0, 0, 0, 0, 0, 0], [-1, 1, 1, 1, 2, 2], [7, 0, 0, 0, 0, -1]):
        with pytest.raises(xt.XtError):
            ctx.set_group_shape(xt.PREC_FP64, bad)
    with pytest.raises(xt.XtError):
        ctx.set_group_shape(5, [1, 1, 1, 1, 1, 1])
    ctx.close()


def test_gather_and_expansion_errors(xt):
    """Host-only contexts and bad task tuples are errors, not crashes, for
    the device gather, the amplitude expansion and the Cholesky upload."""
    L = xt.lib()
    ctx = xt.CCSDTriples(-1)
    ctx.set_tiling(xt.Tiling(8, 24, [4, 4], [12, 12]))
    one = C.c_void_p(8)
    z = np.zeros(9 * 16, dtype=np.int32)
    zp = z.ctypes.data_as(C.POINTER(C.c_int))
    task = np.array([0, 0, 0, 4, 4, 4], dtype=np.int64)
    tp = task.ctypes.data_as(C.POINTER(C.c_long))
    rc = L.xt_gather_task(ctx._ctx, tp, None, *[one] * 6, *[zp] * 7)
    assert rc == xt.XT_ENODEV
    rc = L.xt_gather_task(ctx._ctx, tp, None, *[None] * 6, *[zp] * 7)
    assert rc == xt.XT_EINVAL
    assert L.xt_expand_amplitudes(ctx._ctx, None, one, one) == xt.XT_ENODEV
    sz = ctx.packed_sizes()
    assert sz["max_d1"] == 9 * 2 and sz["max_d2"] == 9 * 2  # alpha tiles per space
    ctx.close()


def test_sum_task_energies_order_and_errors(xt):
    """CPU: xt_sum_task_energies adds in the fixed order of the device
    reduction (256 runs of ceil(n/256) consecutive tasks, then a halving
    tree).  For four tasks that pairs (0, 2) with (1, 3): [1e16, 1, -1e16, 1]
    gives (1e16 - 1e16) + (1 + 1) = 2, where a serial sum gives 1 (1e16 + 1
    rounds to 1e16).  Empty input gives 0; bad arguments are refused."""
    import ctypes as C
    assert xt.sum_task_energies(np.zeros((0, 2))) == (0.0, 0.0)
    p = np.array([[1e16, 1e16], [1.0, 1.0], [-1e16, -1e16], [1.0, 1.0]])
    assert xt.sum_task_energies(p) == (2.0, 2.0)
    assert ((1e16 + 1.0) - 1e16) + 1.0 == 1.0
    r = np.random.default_rng(5).normal(size=(10007, 2))
    e4, e5 = xt.sum_task_energies(r)
    assert abs(e4 - r[:, 0].sum()) < 1e-10 and abs(e5 - r[:, 1].sum()) < 1e-10
    e = np.zeros(2)
    ptr = e.ctypes.data_as(C.POINTER(C.c_double))
    assert xt.lib().xt_sum_task_energies(-1, ptr, ptr) != 0
    assert xt.lib().xt_sum_task_energies(3, None, ptr) != 0


def test_memory_required(xt):
    """CPU: the device-memory figures the uploads check before allocating
    (in place of check_memory_req, hybrid.cpp:19-41) at O=48, V=384 and
    O=64, V=512: resident FP64 operands, upload staging and FP16 split copies
    in GB; the GPU test compares them with the memory an upload really takes.
    No tiling -> refused."""
    ctx = xt.CCSDTriples(-1)
    with pytest.raises(xt.XtError):
        ctx.memory_required()
    for (O, V), want in (((48, 384), (57.8, 3.06, 27.5)), ((64, 512), (182.5, 9.66, 87.0))):
        ctx.set_tiling(xt.Tiling.uneven(O, V, (6, 20), (16, 48), seed=1))
        r = ctx.memory_required()
        got = (r["operands"] / 1e9, r["staging"] / 1e9, r["split"] / 1e9)
        assert np.allclose(got, want, atol=0.05), (O, V, got)
        # staging at the peak is T2 and ovoo: vvvo is freed before xva exists;
        # the Cholesky upload's GEMM workspace (L, its slices, (ai|bj)) stays below that
        t2ov = 8.0 * (V * V * O * O + O * O * O * V)
        assert r["staging"] == t2ov
        assert ctx.memory_required(nchol=4 * (O + V))["staging"] <= r["staging"]
    ctx.close()
