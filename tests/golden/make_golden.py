"""Writes the golden fixtures of the GPU parity tests (runs on the CPU).

Inputs come from the product's seeded generator (xt_synth_closed_shell), so
the GPU tests regenerate the identical arrays on the GPU host; the fixture
stores checksums of every input array to prove that.  Expected energies come
from the CPU oracle (oracle/ccsd_t_oracle.cpp, the restatement of ExaChem's
CPU fused (T) path).

    python tests/golden/make_golden.py small      # config 0 + small uneven cases (~1 min)
    python tests/golden/make_golden.py wide       # adds the wide-tile case to the same file
    python tests/golden/make_golden.py c1sample   # sampled tasks of O=40,V=300 (long)
    python tests/golden/make_golden.py c3sample   # sampled tasks of O=48,V=384 uneven (long)
    python tests/golden/make_golden.py c3wide     # O=48,V=384 with a 20-wide occupied tile (long)
    python tests/golden/make_golden.py c3phys     # c3sample's tasks at a physical scale (long)
    python tests/golden/make_golden.py c4sample   # sampled tasks of O=64,V=512 uneven (long)

The O=48/V=384 and O=64/V=512 samples run the oracle on the generator's
inputs formed element by element where the gathers read them
(oracle.run_generated, pinned bit for bit against the product's generator),
because the O=64/V=512 vvvo block alone (68.7 GB) exceeds host memory.
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import exachem_amd as X  # noqa: E402
from oracle import oracle as O  # noqa: E402

SEED = 20241202
SIGMA, T1_SIGMA = 0.05, 1e-3


def checksums(s):
    return {k: [float(np.sum(s[k])), float(np.sum(s[k] * s[k]))] for k in sorted(s)}


def oracle_case(name, O_, V_, occ, vir, seed, restricted=True, which=None):
    s = X.synth_closed_shell(O_, V_, seed, SIGMA, T1_SIGMA)
    t = O.Tiling(occ, occ, vir, vir)
    evl = np.concatenate([s["eo"], s["eo"], s["ev"], s["ev"]])
    t0 = time.time()
    e4, e5, pt = O.run_spatial(t, evl, s, restricted, which)
    dt = time.time() - t0
    tasks, fac = O.task_list(t, restricted)
    tot, parts, per_task, _ = O.flops(t, restricted, per_task=True)
    print(f"{name}: E4={e4!r} E5={e5!r} ({dt:.1f}s, {O.num_threads()} threads)", flush=True)
    out = dict(name=name, O=O_, V=V_, occ_tiles=occ, vir_tiles=vir, seed=seed, sigma=SIGMA,
               t1_sigma=T1_SIGMA, restricted=restricted, checksums=checksums(s),
               n_tasks=int(len(fac)), total_num_ops=int(tot), oracle_seconds=dt,
               oracle_threads=O.num_threads())
    if which is None:
        out.update(E4=e4, E5=e5, task_partials=pt.tolist())
    else:
        out.update(which=[int(w) for w in which], task_partials=pt.tolist(),
                   task_tuples=tasks[np.asarray(which)].tolist(),
                   task_flops=[float(per_task[w]) for w in which])
    return out


def small():
    cases = [
        oracle_case("config0", 8, 48, [8], [16, 16, 16], SEED),
        # uneven, non-multiple-of-4 tiles; restricted and unrestricted
        oracle_case("uneven_small", 7, 22, [3, 4], [5, 9, 8], SEED + 1),
        oracle_case("uneven_small_unrestricted", 7, 22, [3, 4], [5, 9, 8], SEED + 1,
                    restricted=False),
        oracle_case("one_tile", 5, 13, [5], [13], SEED + 2),
    ]
    with open(os.path.join(HERE, "parity_small.json"), "w") as f:
        json.dump(cases, f, indent=1)


def wide():
    """A one-orbital occupied tile next to a 6-wide one, and a 50-wide virtual
    tile (wider than one 48-element gather slab, seven 8-wide boxes of the
    mixed kernel) next to a 3-wide one.  Added to parity_small.json without
    recomputing the other cases."""
    path = os.path.join(HERE, "parity_small.json")
    cases = [c for c in json.load(open(path)) if c["name"] != "wide_tiles"]
    cases.append(oracle_case("wide_tiles", 7, 53, [6, 1], [50, 3], SEED + 3))
    with open(path, "w") as f:
        json.dump(cases, f, indent=1)


def c1sample():
    occ, vir = X.triples.uniform_tiles(40, 16), X.triples.uniform_tiles(300, 32)
    t = O.Tiling(occ, occ, vir, vir)
    tasks, fac = O.task_list(t)
    tot, parts, per_task, _ = O.flops(t, True, per_task=True)
    rng = np.random.default_rng(7)
    pick = {int(np.argmax(per_task)), int(np.argmin(per_task)), 0, len(fac) - 1}
    # tasks with repeated tiles (factor 1/6 and 1/2 paths) and a mixed-spin task
    for k in range(len(fac)):
        h1, h2, h3, p4, p5, p6 = tasks[k]
        if h1 == h2 == h3 and p4 == p5 == p6:
            pick.add(k)
            break
    pick |= set(int(x) for x in rng.choice(len(fac), 4, replace=False))
    which = sorted(pick)
    # one task at a time so progress is visible; the oracle's per-task values do
    # not depend on which other tasks run in the same call
    parts_ = []
    for w in which:
        c = oracle_case("config1_sample", 40, 300, occ, vir, SEED, which=[w])
        parts_.append(c)
    case = parts_[0]
    case["which"] = which
    case["task_partials"] = [p["task_partials"][0] for p in parts_]
    case["task_tuples"] = [p["task_tuples"][0] for p in parts_]
    case["task_flops"] = [p["task_flops"][0] for p in parts_]
    case["oracle_seconds"] = sum(p["oracle_seconds"] for p in parts_)
    with open(os.path.join(HERE, "parity_c1_sample.json"), "w") as f:
        json.dump(case, f, indent=1)


def pick_tasks(tasks, per_task, tiles, spins, n_quant=(0.10, 0.25, 0.50, 0.75), seed=7):
    """Seeded picks over the flop distribution plus the structural cases:
    the smallest task, one task near each flop quantile, the cheapest task
    holding both the widest occupied and the widest virtual tile, the
    cheapest task with all three hole tiles and all three particle tiles
    equal (the 1/6 factor paths), the cheapest with exactly two equal of each
    (the 1/2 paths), and the cheapest whose hole tiles mix alpha and beta."""
    rng = np.random.default_rng(seed)
    order = np.argsort(per_task, kind="stable")
    n = len(order)
    pick = {"smallest": int(order[0])}
    for q in n_quant:
        lo, hi = int((q - 0.02) * n), int((q + 0.02) * n)
        pick[f"q{int(100 * q)}"] = int(order[rng.integers(lo, hi)])
    dims = tiles[tasks]
    sp = spins[tasks]
    occ_w = dims[:, :3].max(axis=1)
    vir_w = dims[:, 3:].max(axis=1)
    wmax_o = max(int(x) for x in dims[:, :3].reshape(-1))
    wmax_v = max(int(x) for x in dims[:, 3:].reshape(-1))

    def cheapest(mask, name):
        idx = np.nonzero(mask)[0]
        if len(idx):
            pick[name] = int(idx[np.argmin(per_task[idx])])

    cheapest((occ_w == wmax_o) & (vir_w == wmax_v), "widest_tiles")
    t = tasks
    cheapest((t[:, 0] == t[:, 1]) & (t[:, 1] == t[:, 2]) & (t[:, 3] == t[:, 4]) &
             (t[:, 4] == t[:, 5]), "all_equal_tiles")
    two_h = ((t[:, 0] == t[:, 1]) != (t[:, 1] == t[:, 2]))
    two_p = ((t[:, 3] == t[:, 4]) != (t[:, 4] == t[:, 5]))
    cheapest(two_h & two_p, "two_equal_tiles")
    cheapest((sp[:, 0] != sp[:, 2]), "mixed_spin_holes")
    return pick


def large_sample(name, O_, V_, occ, vir, out, sigma=SIGMA, which=None):
    t = O.Tiling(occ, occ, vir, vir)
    tasks, fac = O.task_list(t)
    tot, parts, per_task, _ = O.flops(t, True, per_task=True)
    picks = pick_tasks(tasks, per_task, t.tiles, t.spins)
    if which is None:
        which = sorted(set(picks.values()))
    print(f"{name}: {len(fac)} tasks, {tot:.4e} flops; picks {picks}; "
          f"{sum(per_task[w] for w in which):.3e} flops to run", flush=True)
    rows = []
    for w in which:
        t0 = time.time()
        e4, e5, pt = O.run_generated(t, O_, V_, SEED, sigma, T1_SIGMA, which=[w])
        dt = time.time() - t0
        print(f"  task {w} {tasks[w].tolist()} dims {t.tiles[tasks[w]].tolist()} "
              f"{per_task[w]:.3e} flops: E4 {pt[0, 0]!r} E5 {pt[0, 1]!r} ({dt:.1f}s)", flush=True)
        rows.append((w, pt[0].tolist(), dt))
    case = dict(name=name, O=O_, V=V_, occ_tiles=list(occ), vir_tiles=list(vir), seed=SEED,
                sigma=sigma, t1_sigma=T1_SIGMA, restricted=True, n_tasks=int(len(fac)),
                total_num_ops=int(tot), which=which,
                roles={k: v for k, v in picks.items()},
                task_tuples=[tasks[w].tolist() for w in which],
                task_dims=[t.tiles[tasks[w]].tolist() for w in which],
                task_factors=[float(fac[w]) for w in which],
                task_flops=[float(per_task[w]) for w in which],
                task_partials=[r[1] for r in rows],
                oracle_seconds=sum(r[2] for r in rows), oracle_threads=O.num_threads(),
                inputs="oracle.run_generated (the product generator, restated)")
    with open(os.path.join(HERE, out), "w") as f:
        json.dump(case, f, indent=1)


def c3sample():
    t = X.Tiling.uneven(48, 384, (6, 20), (16, 48), seed=SEED)
    large_sample("c3_O48_V384_uneven", 48, 384, t.occ, t.vir, "parity_c3_sample.json")


def c3phys():
    """The O=48, V=384 sample's tasks at a physical energy scale: integrals
    scaled by 2^-8 (sigma 0.05/256), so that E(T) of the whole workload is
    about 0.2 Eh instead of about 7e8 Eh (E4 scales as sigma^4)."""
    t = X.Tiling.uneven(48, 384, (6, 20), (16, 48), seed=SEED)
    which = json.load(open(os.path.join(HERE, "parity_c3_sample.json")))["which"]
    large_sample("c3_O48_V384_uneven_physical", 48, 384, t.occ, t.vir,
                 "parity_c3_phys_sample.json", sigma=SIGMA / 256, which=which)


def c3wide():
    """O=48, V=384 with a 20-wide occupied tile beside the seeded virtual
    tiles (widest 48): the seeded uneven tiling's widest occupied tile is 15."""
    t = X.Tiling.uneven(48, 384, (6, 20), (16, 48), seed=SEED)
    large_sample("c3_O48_V384_occ20", 48, 384, [20, 6, 14, 8], t.vir,
                 "parity_c3_wide_sample.json")


def c4sample():
    t = X.Tiling.uneven(64, 512, (6, 20), (16, 48), seed=SEED)
    large_sample("c4_O64_V512_uneven", 64, 512, t.occ, t.vir, "parity_c4_sample.json")


if __name__ == "__main__":
    {"small": small, "wide": wide, "c1sample": c1sample, "c3sample": c3sample,
     "c3wide": c3wide, "c3phys": c3phys, "c4sample": c4sample}[sys.argv[1]]()
