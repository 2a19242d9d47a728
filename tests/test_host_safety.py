"""CPU tests of the host-side safety net: the exact fixed-point energy sums
under the address and undefined-behaviour sanitizers, rejection of
non-finite inputs before any device work, and the oracle's restatement of the
input generator (which the O=48/V=384 and O=64/V=512 fixtures are computed
from) against the product's generator."""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _build_and_run(src, tmp_path, extra=()):
    exe = str(tmp_path / os.path.splitext(os.path.basename(src))[0])
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all", *extra, src, "-o", exe], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([exe], capture_output=True, text=True, env=env)
    return r


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_fixed_point_sums_exact_and_refuse_non_finite(tmp_path):
    r = _build_and_run(os.path.join(HERE, "host", "fx_check.cpp"), tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "fx_check: ok" in r.stdout


def test_upload_rejects_non_finite_inputs(xt):
    O_, V_ = 5, 13
    s = xt.synth_closed_shell(O_, V_, 3, 0.05, 1e-3)
    for key, name, idx, bad in [("vvvo", "vvvo[1758]", (2, 1, 0, 3), np.nan),
                                ("t2", "t2[3]", (0, 0, 0, 3), np.inf),
                                ("eo", "eps_occ[2]", (2,), -np.inf),
                                ("ovoo", "ovoo[", (4, 12, 4, 4), np.nan)]:
        ctx = xt.CCSDTriples(-1)  # host-only: the check precedes every device call
        ctx.set_tiling(xt.Tiling(O_, V_, [O_], [V_]))
        bad_s = {k: v.copy() for k, v in s.items()}
        bad_s[key][idx] = bad
        with pytest.raises(xt.XtError) as e:
            ctx.upload(bad_s)
        assert e.value.code == xt.XT_ERANGE, str(e.value)
        assert f"{name}" in str(e.value) and "not finite" in str(e.value), str(e.value)
        ctx.close()
    # finite inputs get past the check to the host-only refusal
    ctx = xt.CCSDTriples(-1)
    ctx.set_tiling(xt.Tiling(O_, V_, [O_], [V_]))
    with pytest.raises(xt.XtError) as e:
        ctx.upload(s)
    assert e.value.code == xt.XT_ENODEV
    with pytest.raises(xt.XtError) as e:
        ctx.synth_upload(1, float("nan"), 1e-3)
    assert e.value.code == xt.XT_ERANGE
    L = np.zeros((O_ + V_, O_ + V_, 4))
    L[1, 2, 3] = np.nan
    with pytest.raises(xt.XtError) as e:
        ctx.upload_cholesky(s["eo"], s["ev"], s["t1"], s["t2"], L)
    assert e.value.code == xt.XT_ERANGE and "chol" in str(e.value)
    ctx.close()


@pytest.mark.parametrize("O_,V_,seed", [(7, 22, 5), (5, 13, 20241202), (12, 31, 9)])
def test_oracle_generator_matches_product(xt, oracle, O_, V_, seed):
    a = xt.synth_closed_shell(O_, V_, seed, 0.05, 1e-3)
    b = oracle.gen_synth(O_, V_, seed, 0.05, 1e-3)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    # a power-of-two sigma ratio scales every integral exactly
    c = oracle.gen_synth(O_, V_, seed, 0.05 / 256, 1e-3)
    for k in ("vvoo", "ovoo", "vvvo", "t2"):
        assert np.array_equal(c[k] * 256.0, a[k]), k


def test_generated_run_equals_array_run(oracle):
    O_, V_, seed = 7, 22, 11
    s = oracle.gen_synth(O_, V_, seed)
    t = oracle.Tiling.closed_shell(O_, V_, [3, 4], [5, 9, 8])
    evl = np.concatenate([s["eo"], s["eo"], s["ev"], s["ev"]])
    e1 = oracle.run_spatial(t, evl, s, True)
    e2 = oracle.run_generated(t, O_, V_, seed)
    assert e1[0] == e2[0] and e1[1] == e2[1]
    assert np.array_equal(e1[2], e2[2])
    which = [3, 0, 7]
    e3 = oracle.run_generated(t, O_, V_, seed, which=which)
    assert np.array_equal(e3[2], e1[2][which])
