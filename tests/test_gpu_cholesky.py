"""V2 blocks built on the GPU from Cholesky vectors (xt_upload_cholesky)
against the integrals formed on the host from the same vectors and uploaded
through xt_upload_closed_shell, whose (T) path the oracle tests pin.

The host integrals are numpy einsums in FP64; the device ones are rocBLAS
DGEMMs over the same Cholesky index, summed in another order, so integrals
agree to ~1e-15 relative and the energies to ~1e-14 relative.  The bar:
FP64 |dE| <= 1e-10 Eh + 1e-12 |E| for E[4], E[5] and every task; mixed
|dE| <= 1e-7 |E4| (the mixed form's own tolerance)."""
import numpy as np
import pytest


def chol_case(O, V, K, seed):
    rng = np.random.default_rng(seed)
    nmo = O + V
    L = rng.normal(0.0, 1.0, (nmo, nmo, K))
    L = 0.5 * (L + L.transpose(1, 0, 2)) * (0.05 / np.sqrt(K)) ** 0.5  # symmetric in (p, q)
    eo = np.sort(rng.uniform(-2.0, -0.3, O))
    ev = np.sort(rng.uniform(0.1, 3.0, V))
    occ, vir = L[:O, :O], L[O:, :O]
    vvoo = np.einsum("aiQ,bjQ->abij", vir, vir)               # (ai|bj)
    vvoo = 0.5 * (vvoo + vvoo.transpose(1, 0, 3, 2))          # exact (ai|bj) = (bj|ai)
    ovoo = np.einsum("iaQ,jkQ->iajk", L[:O, O:], occ)          # (ia|jk)
    vvvo = np.einsum("abQ,ciQ->abci", L[O:, O:], vir)          # (ab|ci)
    D = (eo[None, None, :, None] + eo[None, None, None, :]
         - ev[:, None, None, None] - ev[None, :, None, None])
    t2 = vvoo / D
    t1 = 1e-3 * rng.normal(0.0, 1.0, (V, O))
    return dict(eo=eo, ev=ev, t1=t1, t2=t2, vvoo=vvoo, ovoo=ovoo, vvvo=vvvo), L


def test_cholesky_upload_rejects_host_only(xt):
    """CPU: a host-only context refuses the device construction."""
    s, L = chol_case(4, 12, 8, 1)
    ctx = xt.CCSDTriples(-1)
    ctx.set_tiling(xt.Tiling(4, 12, [4], [12]))
    with pytest.raises(xt.XtError):
        ctx.upload_cholesky(s["eo"], s["ev"], s["t1"], s["t2"], L)
    ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("O,V,K,occ,vir", [(10, 40, 150, (3, 6), (8, 16)),
                                           (8, 48, 100, (8, 8), (16, 16))])
def test_cholesky_v2_matches_host_integrals(xt, O, V, K, occ, vir):
    s, L = chol_case(O, V, K, 2024 + O)
    tiling = xt.Tiling.uneven(O, V, occ, vir, seed=5)
    a = xt.CCSDTriples(0)
    a.set_tiling(tiling)
    a.upload(s)
    b = xt.CCSDTriples(0)
    b.set_tiling(tiling)
    b.upload_cholesky(s["eo"], s["ev"], s["t1"], s["t2"], L)
    # the memory guard's figure covers the Cholesky upload's high point (its
    # own allocations; rocBLAS workspace aside), as for the regular upload
    req, peak = b.memory_required(nchol=K), b.upload_peak()
    print(f"[gpu] cholesky upload peak {peak} B, estimate {req['operands'] + req['staging']} B",
          flush=True)
    assert 0 < peak <= req["operands"] + req["staging"] + 2**20
    for prec in (xt.PREC_FP64, xt.PREC_MIXED):
        ra, rb = a.run(precision=prec), b.run(precision=prec)
        print(f"[gpu] cholesky O={O} V={V} K={K} prec {prec}: E4 {ra['e4']!r} / {rb['e4']!r} "
              f"E5 {ra['e5']!r} / {rb['e5']!r}", flush=True)
        for ka in ("e4", "e5"):
            d = abs(ra[ka] - rb[ka])
            if prec == xt.PREC_FP64:
                assert d <= 1e-10 + 1e-12 * abs(ra[ka]), (ka, d)
            else:
                assert d <= 1e-7 * abs(ra["e4"]), (ka, d)
        if prec == xt.PREC_FP64:
            pa, pb = ra["task_partials"], rb["task_partials"]
            assert np.all(np.abs(pa - pb) <= 1e-10 + 1e-12 * np.abs(pa))
    a.close()
    b.close()
