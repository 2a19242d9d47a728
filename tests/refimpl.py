"""Independent numpy formulations used as checks of the oracle (test-only).

* ``synth_spatial``: small seeded closed-shell spatial inputs with the exact
  permutational symmetries the generator promises (pure numpy, independent of
  the product's C++ generator).
* ``bruteforce_so``: dense, tile-free spin-orbital evaluation of
      t3d = -A_h A_p [sum_h7 T2(p4,p5,h1,h7) <h2 h3||h7 p6>]
            -A_h A_p [sum_p7 T2(p4,p7,h1,h2) <h3 p7||p5 p6>]
      t3s =  A_h A_p [T1(p4,h1) <h3 h2||p6 p5>]
  over the 9 coset representatives of each antisymmetriser (signs =
  permutation parities), E4 = 1/36 sum t3d^2/D, E5 = 1/36 sum t3d(t3d+t3s)/D.
  It uses no tiles, gathers, exec slots or factors.
* ``spatial_closed_shell``: textbook closed-shell (T) (W/V intermediates with
  the 4W+W(bca)+W(cab) weights) on spatial orbitals -- a derivation that shares
  nothing with the spin-orbital path.
"""
import itertools

import numpy as np


def synth_spatial(O, V, seed, sigma=0.05, t1_sigma=1e-3):
    rng = np.random.default_rng(seed)
    eo = np.sort(rng.uniform(-2.0, -0.3, O))
    ev = np.sort(rng.uniform(0.1, 3.0, V))
    x = rng.normal(0, sigma, (V, V, O, O))
    # (ai|bj) = (bj|ai): vvoo[a,b,i,j] = vvoo[b,a,j,i] by mirroring
    vvoo = x.copy()
    for a in range(V):
        for b in range(V):
            for i in range(O):
                for j in range(O):
                    if (b, j) < (a, i):
                        vvoo[a, b, i, j] = x[b, a, j, i]
    y = rng.normal(0, sigma, (O, V, O, O))
    ovoo = y.copy()
    for j in range(O):
        for k in range(O):
            if k < j:
                ovoo[:, :, j, k] = y[:, :, k, j]
    z = rng.normal(0, sigma, (V, V, V, O))
    vvvo = z.copy()
    for a in range(V):
        for b in range(V):
            if b < a:
                vvvo[a, b] = z[b, a]
    D = (eo[None, None, :, None] + eo[None, None, None, :]) - (ev[:, None, None, None] + ev[None, :, None, None])
    t2 = vvoo / D
    t1 = rng.normal(0, t1_sigma, (V, O))
    return dict(eo=eo, ev=ev, t1=t1, t2=t2, vvoo=vvoo, ovoo=ovoo, vvvo=vvvo)


def _perm_sign(p):
    s = 1
    p = list(p)
    for i in range(len(p)):
        for j in range(i + 1, len(p)):
            if p[i] > p[j]:
                s = -s
    return s


def bruteforce_so(so, eps_occ_so, eps_vir_so):
    t1, t2, vijab, vijka, viabc = so
    # X1[h1,h2,h3,p4,p5,p6] = sum_h7 T2(p4,p5,h1,h7) vijka(h2,h3,h7,p6)
    X1 = np.einsum('abim,jkmc->ijkabc', t2, vijka, optimize=True)
    # X2[h1,h2,h3,p4,p5,p6] = sum_p7 T2(p4,p7,h1,h2) viabc(h3,p7,p5,p6)
    X2 = np.einsum('aeij,kebc->ijkabc', t2, viabc, optimize=True)
    # S[h1,h2,h3,p4,p5,p6] = T1(p4,h1) vijab(h3,h2,p6,p5)
    S = np.einsum('ai,kjcb->ijkabc', t1, vijab, optimize=True)
    hp = [(0, 1, 2), (1, 0, 2), (2, 0, 1)]
    pp = [(0, 1, 2), (1, 2, 0), (0, 2, 1)]
    t3d = np.zeros_like(X1)
    t3s = np.zeros_like(X1)
    for h in hp:
        for p in pp:
            sgn = _perm_sign(h) * _perm_sign(p)
            # Y[h1,h2,h3,p4,p5,p6] = X[h_{h[0]},h_{h[1]},h_{h[2]},p_{p[0]},p_{p[1]},p_{p[2]}]
            src = list(h) + [3 + q for q in p]
            t3d += -sgn * np.einsum(X1, src, [0, 1, 2, 3, 4, 5])
    # d2 uses its own coset representatives: h pairs (h1,h2),(h2,h3),(h1,h3) with
    # the single h in the v2 slot; p single in the T2 slot.
    hp2 = [(0, 1, 2), (1, 2, 0), (0, 2, 1)]
    pp2 = [(0, 1, 2), (1, 0, 2), (2, 0, 1)]
    for h in hp2:
        for p in pp2:
            sgn = _perm_sign(h) * _perm_sign(p)
            src = list(h) + [3 + q for q in p]
            t3d += -sgn * np.einsum(X2, src, [0, 1, 2, 3, 4, 5])
    hps = [(0, 1, 2), (1, 0, 2), (2, 0, 1)]
    pps = [(0, 1, 2), (1, 0, 2), (2, 0, 1)]
    for h in hps:
        for p in pps:
            sgn = _perm_sign(h) * _perm_sign(p)
            src = list(h) + [3 + q for q in p]
            t3s += sgn * np.einsum(S, src, [0, 1, 2, 3, 4, 5])
    eo, ev = eps_occ_so, eps_vir_so
    D = (eo[:, None, None, None, None, None] + eo[None, :, None, None, None, None] + eo[None, None, :, None, None, None]
         - ev[None, None, None, :, None, None] - ev[None, None, None, None, :, None] - ev[None, None, None, None, None, :])
    E4 = np.sum(t3d * t3d / D) / 36.0
    E5 = np.sum(t3d * (t3d + t3s) / D) / 36.0
    return E4, E5


def spatial_closed_shell(s):
    """Closed-shell (T) on spatial orbitals; returns (E[T], E(T))."""
    eo, ev, t1, t2 = s['eo'], s['ev'], s['t1'], s['t2']
    vvoo, ovoo, vvvo = s['vvoo'], s['ovoo'], s['vvvo']
    O, V = len(eo), len(ev)
    # Mulliken integrals: (bd|ai) = vvvo[b,d,a,i]; (ck|jl) = ovoo[k,c,j,l]; (ai|bj) = vvoo[a,b,i,j]
    # W_ijk^abc = P[ sum_d (bd|ai) t_kj^cd - sum_l (ck|jl) t_il^ab ]
    # t_ij^ab = t2[a,b,i,j]
    base = (np.einsum('bdai,cdkj->ijkabc', vvvo, t2, optimize=True)
            - np.einsum('kcjl,abil->ijkabc', ovoo, t2, optimize=True))
    W = np.zeros_like(base)
    # simultaneous permutations of (i,a),(j,b),(k,c)
    for p in itertools.permutations(range(3)):
        src = [p[0], p[1], p[2], 3 + p[0], 3 + p[1], 3 + p[2]]
        W += np.einsum(base, src, [0, 1, 2, 3, 4, 5])
    Vd = (np.einsum('ai,bcjk->ijkabc', t1, vvoo, optimize=True)
          + np.einsum('bj,acik->ijkabc', t1, vvoo, optimize=True)
          + np.einsum('ck,abij->ijkabc', t1, vvoo, optimize=True))
    Vt = W + Vd
    D = (eo[:, None, None, None, None, None] + eo[None, :, None, None, None, None] + eo[None, None, :, None, None, None]
         - ev[None, None, None, :, None, None] - ev[None, None, None, None, :, None] - ev[None, None, None, None, None, :])
    def tr(X, order):  # X^{order(abc)}_{ijk}
        return np.einsum(X, [0, 1, 2] + [3 + o for o in order], [0, 1, 2, 3, 4, 5])
    Y = 4 * W + tr(W, (1, 2, 0)) + tr(W, (2, 0, 1))
    E4 = np.sum(Y * (W - tr(W, (2, 1, 0))) / D) / 3.0
    E5 = np.sum(Y * (Vt - tr(Vt, (2, 1, 0))) / D) / 3.0
    return E4, E5


def so_energies(s, tiling=None):
    O, V = len(s['eo']), len(s['ev'])
    eo = np.concatenate([s['eo'], s['eo']])
    ev = np.concatenate([s['ev'], s['ev']])
    return eo, ev
