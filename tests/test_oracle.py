"""Pins the CPU oracle (the checker) -- see the header of oracle/oracle.py.

* exact known answers: total_num_ops of the reference's CI outputs;
* mathematics: dense tile-free brute force, an independent spatial
  closed-shell (T) formula, tile invariance, restricted == unrestricted,
  T1 = 0 => E5 = E4, scaling laws, spin-orbital antisymmetry.
Each check is chosen so that a plausible restatement mistake (wrong sign of
one equation, transposed operand, wrong exec slot, wrong factor) fails it;
test_planted_mistakes_are_caught demonstrates that on the oracle itself.
"""
import json
import os

import numpy as np
import pytest

from tests import refimpl as R

HERE = os.path.dirname(os.path.abspath(__file__))


def _tiling(O_, case):
    if "Oa" in case:
        T = O_.Tiling
        return T([case["Oa"]] if case["Oa"] <= case["ts"] else T.uniform(case["Oa"], case["ts"]),
                 T.uniform(case["Ob"], case["ts"]), T.uniform(case["Va"], case["ts"]),
                 T.uniform(case["Vb"], case["ts"]))
    return O_.Tiling.closed_shell(case["O"], case["V"], case["ts"], case["ts"])


@pytest.mark.parametrize("idx", range(5))
def test_total_num_ops_known_answers(oracle, idx):
    cases = json.load(open(os.path.join(HERE, "golden", "kat_total_num_ops.json")))["cases"]
    case = cases[idx]
    total, parts = oracle.flops(_tiling(oracle, case), case["restricted"])
    assert total == case["total_num_ops"], (case["name"], total, parts)


def test_task_counts_of_configs(oracle):
    # config 0 (O=8,V=48, tiles 8/16) and config 1 (O=40,V=300, tiles 16/32)
    t0 = oracle.Tiling.closed_shell(8, 48, 8, 16)
    t1 = oracle.Tiling.closed_shell(40, 300, 16, 32)
    assert len(oracle.task_list(t0)[1]) == 28
    assert oracle.flops(t0)[0] == 55717134336
    assert len(oracle.task_list(t1)[1]) == 12100
    assert oracle.flops(t1)[0] == 4459374691614720


def _so(oracle, s):
    O_, V_ = len(s["eo"]), len(s["ev"])
    return oracle.expand_closed_shell(O_, V_, s["t1"], s["t2"], s["vvoo"], s["ovoo"], s["vvvo"])


def _evl(s):
    return np.concatenate([s["eo"], s["eo"], s["ev"], s["ev"]])


def test_spin_orbital_antisymmetry(oracle):
    s = R.synth_spatial(3, 5, 11)
    t1, t2, ijab, ijka, iabc = _so(oracle, s)
    assert np.array_equal(t2, -t2.transpose(1, 0, 2, 3))
    assert np.array_equal(t2, -t2.transpose(0, 1, 3, 2))
    assert np.array_equal(ijab, -ijab.transpose(1, 0, 2, 3))
    assert np.array_equal(ijab, -ijab.transpose(0, 1, 3, 2))
    assert np.array_equal(ijka, -ijka.transpose(1, 0, 2, 3))
    assert np.array_equal(iabc, -iabc.transpose(0, 1, 3, 2))


@pytest.mark.parametrize("O_,V_,seed", [(3, 5, 1), (4, 6, 2), (2, 7, 3)])
def test_oracle_matches_bruteforce_and_spatial_formula(oracle, O_, V_, seed):
    s = R.synth_spatial(O_, V_, seed)
    so = _so(oracle, s)
    eo, ev = R.so_energies(s)
    bf4, bf5 = R.bruteforce_so(so, eo, ev)
    sp4, sp5 = R.spatial_closed_shell(s)
    scale = abs(bf4)
    assert abs(bf4 - sp4) <= 1e-12 * scale and abs(bf5 - sp5) <= 1e-12 * scale
    for occ_ts, vir_ts in [(O_, V_), (2, 3), (1, 2)]:
        t = oracle.Tiling.closed_shell(O_, V_, min(occ_ts, O_), min(vir_ts, V_))
        for restricted in (True, False):
            e4, e5, _ = oracle.run(t, _evl(s), so, restricted)
            assert abs(e4 - bf4) <= 1e-12 * scale, (occ_ts, vir_ts, restricted)
            assert abs(e5 - bf5) <= 1e-12 * scale, (occ_ts, vir_ts, restricted)


def test_on_the_fly_elements_equal_dense(oracle):
    s = R.synth_spatial(3, 7, 5)
    t = oracle.Tiling.closed_shell(3, 7, 2, 3)
    a = oracle.run(t, _evl(s), _so(oracle, s))
    b = oracle.run_spatial(t, _evl(s), s)
    assert a[0] == b[0] and a[1] == b[1] and np.array_equal(a[2], b[2])


def test_special_cases(oracle):
    s = R.synth_spatial(3, 6, 9)
    t = oracle.Tiling.closed_shell(3, 6, 2, 4)
    e4, e5, _ = oracle.run_spatial(t, _evl(s), s)
    # T1 = 0  =>  E5 = E4
    z = dict(s, t1=np.zeros_like(s["t1"]))
    z4, z5, _ = oracle.run_spatial(t, _evl(s), z)
    assert z4 == e4 and abs(z5 - z4) <= 1e-15 * abs(z4)
    # E4 ~ lambda^2 under T2 -> lambda T2; E5 - E4 linear in T1
    l4, _, _ = oracle.run_spatial(t, _evl(s), dict(s, t2=3.0 * s["t2"]))
    assert abs(l4 - 9.0 * e4) <= 1e-12 * abs(l4)
    _, m5, _ = oracle.run_spatial(t, _evl(s), dict(s, t1=2.0 * s["t1"]))
    assert abs((m5 - e4) - 2.0 * (e5 - e4)) <= 1e-9 * abs(e5 - e4)
    # all integrals -> mu V with T2 fixed: E4 ~ mu^2
    mu = dict(s, vvoo=2 * s["vvoo"], ovoo=2 * s["ovoo"], vvvo=2 * s["vvvo"])
    m4, _, _ = oracle.run_spatial(t, _evl(s), mu)
    assert abs(m4 - 4.0 * e4) <= 1e-12 * abs(m4)
    # V2 = 0 => 0
    zero = dict(s, vvoo=0 * s["vvoo"], ovoo=0 * s["ovoo"], vvvo=0 * s["vvvo"])
    assert oracle.run_spatial(t, _evl(s), zero)[:2] == (0.0, 0.0)


def test_pauli_forbidden_tasks_vanish(oracle):
    """Two holes (or two particles) in the same one-orbital tile: t3 vanishes
    identically, so the task's E[4] and E[5] are 0 up to FP64 rounding of the
    cancelling terms, squared (exactly 0 for the holes in this evaluation
    order, ~1e-36 of the largest task for the particles).  Each such element
    is a sum of d1/d2/s1 terms that cancel in pairs exchanging the two
    indices, so a wrong sign or slot on one equation leaves it at the scale
    of an ordinary task; every other task is at least 1e-6 of the largest."""
    s = R.synth_spatial(4, 9, 5)
    occ, vir = [3, 1], [7, 1, 1]
    t = oracle.Tiling(occ, occ, vir, vir)
    tasks, _ = oracle.task_list(t)
    _, _, pt = oracle.run_spatial(t, _evl(s), s)
    tiles = np.asarray(t.tiles)

    def repeated_single(ix):
        return any(ix[i] == ix[i + 1] and tiles[ix[i]] == 1 for i in range(2))

    forbidden = np.array([repeated_single(k[:3]) or repeated_single(k[3:]) for k in tasks])
    assert forbidden.sum() >= 4 and (~forbidden).sum() >= 4
    big = np.max(np.abs(pt))
    assert np.all(np.abs(pt[forbidden]) <= 1e-28 * big)
    assert np.all(np.abs(pt[~forbidden]) >= 1e-6 * big)
    holes = np.array([k[0] == k[1] == 1 or k[1] == k[2] == 1 for k in tasks])
    assert holes.sum() >= 4 and np.all(pt[holes] == 0.0)


def test_mp4_triples_identity(oracle):
    """With first-order T2 (t = (ia|jb)/D) and T1 = 0, E[4] is the MP4 triples
    energy; checked against the spatial formula with T1 = 0 (no shared code)."""
    s = R.synth_spatial(3, 5, 21)
    s = dict(s, t1=np.zeros_like(s["t1"]))
    t = oracle.Tiling.closed_shell(3, 5, 3, 5)
    e4, _, _ = oracle.run_spatial(t, _evl(s), s)
    sp4, _ = R.spatial_closed_shell(s)
    assert abs(e4 - sp4) <= 1e-12 * abs(sp4)


def test_planted_mistakes_are_caught(oracle):
    """A wrong sign on a single d1 element, a transposed T2 operand or a wrong
    factor each move the energy far beyond the parity tolerance."""
    s = R.synth_spatial(3, 5, 1)
    so = list(_so(oracle, s))
    eo, ev = R.so_energies(s)
    ref4, _ = R.bruteforce_so(so, eo, ev)
    t = oracle.Tiling.closed_shell(3, 5, 2, 3)
    # transposed T2 operand (a,b,i,j) -> (b,a,i,j) breaks antisymmetry consistency
    bad = list(so)
    bad[1] = so[1].transpose(0, 1, 3, 2).copy() * 0.5 + so[1] * 0.5
    e4b, _, _ = oracle.run(t, _evl(s), bad)
    assert abs(e4b - ref4) > 1e-6 * abs(ref4)
    # one v2ijka element sign-flipped
    bad = list(so)
    bad[3] = so[3].copy()
    bad[3][0, 1, 0, 0] *= -1
    bad[3][1, 0, 0, 0] *= -1
    e4c, _, _ = oracle.run(t, _evl(s), bad)
    assert abs(e4c - ref4) > 1e-9 * abs(ref4)


def test_golden_small_reproduces(oracle, xt):
    """The committed small fixtures are what the oracle produces now."""
    cases = json.load(open(os.path.join(HERE, "golden", "parity_small.json")))
    for case in cases:
        if case["name"] == "config0":
            continue  # 17 s; covered by the GPU parity test against the same fixture
        s = xt.synth_closed_shell(case["O"], case["V"], case["seed"], case["sigma"],
                                  case["t1_sigma"])
        for k, (a, b) in case["checksums"].items():
            assert float(np.sum(s[k])) == a and float(np.sum(s[k] * s[k])) == b, k
        t = oracle.Tiling(case["occ_tiles"], case["occ_tiles"], case["vir_tiles"],
                          case["vir_tiles"])
        e4, e5, pt = oracle.run_spatial(t, _evl(s), s, case["restricted"])
        assert e4 == case["E4"] and e5 == case["E5"]
        assert np.array_equal(pt, np.asarray(case["task_partials"]))
