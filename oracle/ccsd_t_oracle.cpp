// ============================================================================
// TEST INFRASTRUCTURE ONLY.  CPU restatement of ExaChem's closed/open-shell
// fused (T) path, used as the checker for the HIP implementation.  Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// this library; the product path never links or calls it.
//
// It follows, function by function:
//   exachem/cc/ccsd_t/ccsd_t_fused_driver.hpp:349-483   task loop, spin filter, factor
//   exachem/cc/ccsd_t/ccsd_t_all_fused_singles.hpp:34-391   s1 gather
//   exachem/cc/ccsd_t/ccsd_t_all_fused_doubles1.hpp:37-307  d1 gather
//   exachem/cc/ccsd_t/ccsd_t_all_fused_doubles2.hpp:37-360  d2 gather
//   exachem/cc/ccsd_t/ccsd_t_all_fused_cpu.hpp:20-631       per-task CPU kernel + energy
//   exachem/cc/ccsd_t/fused_common.hpp:34-261               analytic flop counter
//   exachem/cc/ccsd_t/*_info_only (singles.hpp:394-553, doubles1.hpp:309-488,
//                                  doubles2.hpp:362-545)
//   exachem/cholesky/v2tensors.cpp:67-90 and cc/ccsd/cd_ccsd_cs_ann.cpp:932-976
//     spin-orbital expansion of closed-shell inputs (oracle_expand_closed_shell)
//
// Two substitutions for libraries absent here (TAMM, hptt):
//   Tensor<T>::get(blockid, buf)  -> row-major copy of the block out of a dense
//                                    spin-orbital array (block_get below)
//   hptt::create_plan(perm, n, alpha, A, size, ..., B, ..., useRowMajor=true)
//                                 -> B[i_0..i_{n-1}] = alpha * A[j], j[perm[k]] = i_k,
//                                    B row-major with dims size[perm[k]] (hptt_permute)
// The LRU caches of the reference are pure memoisation and are omitted.
//
// Parity status of each exported function (how it is pinned) is listed in the
// header comment of oracle/oracle.py.
// ============================================================================
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstddef>
#include <cstring>
#include <tuple>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using Index = size_t;
using Tup6  = std::tuple<Index, Index, Index, Index, Index, Index>;

// Counter-based generator of the synthetic closed-shell inputs, restated from
// the product's host generator so that inputs too large for host memory
// (O=64, V=512: vvvo alone is 68.7 GB) can be formed element by element.
// The two are compared array for array on small sizes (tests/test_oracle.py),
// and the GPU tests that upload the product's arrays reproduce energies
// computed here from this one.
inline uint64_t gen_mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}
inline uint64_t gen_hash3(uint64_t seed, uint64_t stream, uint64_t idx) {
  return gen_mix64(gen_mix64(seed ^ (stream * 0xD6E8FEB86659FD93ULL)) ^ gen_mix64(idx));
}
inline double gen_u01(uint64_t h) {
  return ((double) (h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
inline double gen_gauss(uint64_t seed, uint64_t stream, uint64_t idx) {
  const double u1 = gen_u01(gen_hash3(seed, stream, 2 * idx));
  const double u2 = gen_u01(gen_hash3(seed, stream, 2 * idx + 1));
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586476925286766559 * u2);
}
enum GenStream : uint64_t { G_EO = 1, G_EV = 2, G_VVOO = 3, G_OVOO = 4, G_VVVO = 5, G_T1 = 6 };

struct GenParams {
  long                O = 0, V = 0;
  uint64_t            seed  = 0;
  double              sigma = 0.0, t1s = 0.0;
  std::vector<double> eo, ev;  // sorted orbital energies
  GenParams(long O_, long V_, uint64_t seed_, double sigma_, double t1s_)
    : O(O_), V(V_), seed(seed_), sigma(sigma_), t1s(t1s_), eo(O_), ev(V_) {
    for(long i = 0; i < O; i++) eo[i] = -2.0 + 1.7 * gen_u01(gen_hash3(seed, G_EO, i));
    for(long a = 0; a < V; a++) ev[a] = 0.1 + 2.9 * gen_u01(gen_hash3(seed, G_EV, a));
    std::sort(eo.begin(), eo.end());
    std::sort(ev.begin(), ev.end());
  }
  double t1(long a, long i) const { return t1s * gen_gauss(seed, G_T1, a * O + i); }
  // (ai|bj), drawn at the lexicographically smaller of (a,i,b,j), (b,j,a,i)
  double vvoo(long a, long b, long i, long j) const {
    const bool sw = (b < a) || (b == a && j < i);
    const long a_ = sw ? b : a, i_ = sw ? j : i, b_ = sw ? a : b, j_ = sw ? i : j;
    return sigma * gen_gauss(seed, G_VVOO, ((a_ * O + i_) * V + b_) * O + j_);
  }
  double t2(long a, long b, long i, long j) const {
    return vvoo(a, b, i, j) / ((eo[i] + eo[j]) - (ev[a] + ev[b]));
  }
  double ovoo(long i, long a, long j, long k) const {  // (ia|jk), symmetric in j,k
    return sigma * gen_gauss(seed, G_OVOO, ((i * V + a) * O + std::min(j, k)) * O + std::max(j, k));
  }
  double vvvo(long a, long b, long c, long i) const {  // (ab|ci), symmetric in a,b
    return sigma * gen_gauss(seed, G_VVVO, ((std::min(a, b) * V + std::max(a, b)) * V + c) * O + i);
  }
};

// Closed-shell spatial inputs from which spin-orbital elements are formed
// (oracle_expand_closed_shell documents the formulas): arrays, or the
// generator when gen is set.
struct SpatialSrc {
  long             O = 0, V = 0;
  const double*    t1s  = nullptr;  // [V][O]
  const double*    t2s  = nullptr;  // [V][V][O][O]
  const double*    vvoo = nullptr;  // [V][V][O][O] (ai|bj)
  const double*    ovoo = nullptr;  // [O][V][O][O] (ia|jk)
  const double*    vvvo = nullptr;  // [V][V][V][O] (ab|ci)
  const GenParams* gen  = nullptr;
  double T1(long a, long i) const { return gen ? gen->t1(a, i) : t1s[a * O + i]; }
  double T2(long a, long b, long i, long j) const {
    return gen ? gen->t2(a, b, i, j) : t2s[((a * V + b) * O + i) * O + j];
  }
  double VVOO(long a, long b, long i, long j) const {
    return gen ? gen->vvoo(a, b, i, j) : vvoo[((a * V + b) * O + i) * O + j];
  }
  double OVOO(long i, long a, long j, long k) const {
    return gen ? gen->ovoo(i, a, j, k) : ovoo[((i * V + a) * O + j) * O + k];
  }
  double VVVO(long a, long b, long c, long i) const {
    return gen ? gen->vvvo(a, b, c, i) : vvvo[((a * V + b) * V + c) * O + i];
  }
};

enum Kind { K_T1 = 0, K_T2 = 1, K_IJAB = 2, K_IJKA = 3, K_IABC = 4 };

// One spin-orbital element; spin-orbital order [alpha block, beta block] per space.
inline double so_elem(const SpatialSrc& S, int kind, const long* q) {
  const long O = S.O, V = S.V;
  auto occ = [&](long x, long* sp) { *sp = x / O; return x % O; };
  auto vir = [&](long x, long* sp) { *sp = x / V; return x % V; };
  auto comb = [](bool c1, bool c2, double x, double y) {
    return c1 && c2 ? x - y : (c1 ? x : (c2 ? -y : 0.0));
  };
  switch(kind) {
    case K_T1: {  // t1(a,i)
      long sa, si, A = vir(q[0], &sa), I = occ(q[1], &si);
      return sa == si ? S.T1(A, I) : 0.0;
    }
    case K_T2: {  // t2(a,b,i,j) = [sa=si][sb=sj] t_ij^ab - [sa=sj][sb=si] t_ji^ab
      long sa, sb, si, sj, A = vir(q[0], &sa), B = vir(q[1], &sb), I = occ(q[2], &si),
                           J = occ(q[3], &sj);
      bool c1 = sa == si && sb == sj, c2 = sa == sj && sb == si;
      return comb(c1, c2, c1 ? S.T2(A, B, I, J) : 0.0, c2 ? S.T2(A, B, J, I) : 0.0);
    }
    case K_IJAB: {  // <ij||ab> = (ia|jb) - (ib|ja), (ia|jb) = (ai|bj)
      long si, sj, sa, sb, I = occ(q[0], &si), J = occ(q[1], &sj), A = vir(q[2], &sa),
                           B = vir(q[3], &sb);
      bool c1 = si == sa && sj == sb, c2 = si == sb && sj == sa;
      return comb(c1, c2, c1 ? S.VVOO(A, B, I, J) : 0.0, c2 ? S.VVOO(B, A, I, J) : 0.0);
    }
    case K_IJKA: {  // <ij||ka> = (ik|ja) - (jk|ia), (ik|ja) = (ja|ik)
      long si, sj, sk, sa, I = occ(q[0], &si), J = occ(q[1], &sj), K = occ(q[2], &sk),
                           A = vir(q[3], &sa);
      bool c1 = si == sk && sj == sa, c2 = sj == sk && si == sa;
      return comb(c1, c2, c1 ? S.OVOO(J, A, I, K) : 0.0, c2 ? S.OVOO(I, A, J, K) : 0.0);
    }
    default: {  // K_IABC: <ia||bc> = (ib|ac) - (ic|ab), (ib|ac) = (ac|bi)
      long si, sa, sb, sc, I = occ(q[0], &si), A = vir(q[1], &sa), B = vir(q[2], &sb),
                           C = vir(q[3], &sc);
      bool c1 = si == sb && sa == sc, c2 = si == sc && sa == sb;
      return comb(c1, c2, c1 ? S.VVVO(A, C, B, I) : 0.0, c2 ? S.VVVO(A, B, C, I) : 0.0);
    }
  }
}

// Dense spin-orbital tensors standing in for the TAMM tensors d_t1 (V,O),
// d_t2 (V,V,O,O), v2ijab (O,O,V,V), v2ijka (O,O,O,V), v2iabc (O,V,V,V).
struct SOProblem {
  Index                noab = 0, nvab = 0;
  std::vector<size_t>  k_range, k_offset;
  std::vector<int>     k_spin;
  std::vector<double>  k_evl_sorted;
  size_t               nO = 0, nV = 0;
  const double*        t1     = nullptr;
  const double*        t2     = nullptr;
  const double*        v2ijab = nullptr;
  const double*        v2ijka = nullptr;
  const double*        v2iabc = nullptr;
  const SpatialSrc*    src    = nullptr;  // if set, elements are formed from spatial inputs
};

// Start of tile b inside its own index space (occ tiles start at 0, virt tiles
// start at nO in the MO space, so subtract nO for them).
inline size_t tile_start(const SOProblem& P, Index b) {
  return b < P.noab ? P.k_offset[b] : P.k_offset[b] - P.nO;
}

// Tensor<T>::get(blockid, buf) restated: row-major copy of one block.
// is_virt[d] says whether dimension d of the tensor is a virtual index.
void block_get(const SOProblem& P, int kind, const double* T, const bool* is_virt,
               const Index* mo_tiles, int ndim, std::vector<double>& out) {
  size_t full[4], start[4], size[4];
  for(int d = 0; d < ndim; d++) {
    full[d]  = is_virt[d] ? P.nV : P.nO;
    start[d] = tile_start(P, mo_tiles[d]);
    size[d]  = P.k_range[mo_tiles[d]];
  }
  size_t total = 1;
  for(int d = 0; d < ndim; d++) total *= size[d];
  out.assign(total, 0.0);
  size_t idx[4] = {0, 0, 0, 0};
  for(size_t e = 0; e < total; e++) {
    size_t r = e;
    for(int d = ndim - 1; d >= 0; d--) {
      idx[d] = r % size[d];
      r /= size[d];
    }
    if(P.src) {
      long q[4];
      for(int d = 0; d < ndim; d++) q[d] = (long) (start[d] + idx[d]);
      out[e] = so_elem(*P.src, kind, q);
    }
    else {
      size_t g = 0;
      for(int d = 0; d < ndim; d++) g = g * full[d] + (start[d] + idx[d]);
      out[e] = T[g];
    }
  }
}

// hptt::create_plan(perm, n, alpha, A, sizeA, ..., B, ..., useRowMajor=true) restated.
void hptt_permute(const int* perm, int n, double alpha, const double* A, const int* sizeA,
                  double* B) {
  int    sizeB[4];
  size_t total = 1;
  for(int k = 0; k < n; k++) {
    sizeB[k] = sizeA[perm[k]];
    total *= (size_t) sizeA[k];
  }
  int i[4] = {0, 0, 0, 0}, j[4] = {0, 0, 0, 0};
  for(size_t e = 0; e < total; e++) {
    size_t r = e;
    for(int k = n - 1; k >= 0; k--) {
      i[k] = (int) (r % sizeB[k]);
      r /= sizeB[k];
    }
    for(int k = 0; k < n; k++) j[perm[k]] = i[k];
    size_t a = 0;
    for(int k = 0; k < n; k++) a = a * sizeA[k] + j[k];
    B[e] = alpha * A[a];
  }
}

inline Tup6 zero6() { return std::make_tuple(0, 0, 0, 0, 0, 0); }

// --------------------------------------------------------------------------
// s1 gather: ccsd_t_all_fused_singles.hpp:34-391
// --------------------------------------------------------------------------
void data_s1_new(bool is_restricted, const SOProblem& P, size_t t_h1b, size_t t_h2b, size_t t_h3b,
                 size_t t_p4b, size_t t_p5b, size_t t_p6b, size_t size_T_s1_t1,
                 size_t size_T_s1_v2, int* df_simple_s1_size, int* df_simple_s1_exec,
                 double* df_T_s1_t1, double* df_T_s1_v2, int* df_num_s1_enabled,
                 bool data, size_t* comm_elems) {
  const Index noab = P.noab;
  const auto& k_spin = P.k_spin;
  const auto& k_range = P.k_range;
  Tup6 a3_s1[] = {std::make_tuple(t_p4b, t_p5b, t_p6b, t_h1b, t_h2b, t_h3b),
                  std::make_tuple(t_p4b, t_p5b, t_p6b, t_h2b, t_h1b, t_h3b),
                  std::make_tuple(t_p4b, t_p5b, t_p6b, t_h3b, t_h1b, t_h2b),
                  std::make_tuple(t_p5b, t_p4b, t_p6b, t_h1b, t_h2b, t_h3b),
                  std::make_tuple(t_p5b, t_p4b, t_p6b, t_h2b, t_h1b, t_h3b),
                  std::make_tuple(t_p5b, t_p4b, t_p6b, t_h3b, t_h1b, t_h2b),
                  std::make_tuple(t_p6b, t_p4b, t_p5b, t_h1b, t_h2b, t_h3b),
                  std::make_tuple(t_p6b, t_p4b, t_p5b, t_h2b, t_h1b, t_h3b),
                  std::make_tuple(t_p6b, t_p4b, t_p5b, t_h3b, t_h1b, t_h2b)};
  for(int ia6 = 0; ia6 < 8; ia6++) {
    if(std::get<0>(a3_s1[ia6]) != 0) {
      for(int ja6 = ia6 + 1; ja6 < 9; ja6++)
        if(a3_s1[ia6] == a3_s1[ja6]) a3_s1[ja6] = zero6();
    }
  }
  const size_t s1_max_dima = size_T_s1_t1 / 9;
  const size_t s1_max_dimb = size_T_s1_v2 / 9;

  df_simple_s1_size[0] = (int) k_range[t_h1b];
  df_simple_s1_size[1] = (int) k_range[t_h2b];
  df_simple_s1_size[2] = (int) k_range[t_h3b];
  df_simple_s1_size[3] = (int) k_range[t_p4b];
  df_simple_s1_size[4] = (int) k_range[t_p5b];
  df_simple_s1_size[5] = (int) k_range[t_p6b];

  bool ia6_enabled[9] = {false};
  for(int ia6 = 0; ia6 < 9; ia6++) {
    auto [p4b, p5b, p6b, h1b, h2b, h3b] = a3_s1[ia6];
    if(!((p5b <= p6b) && (h2b <= h3b) && p4b != 0)) continue;
    if(is_restricted &&
       !(k_spin[p4b] + k_spin[p5b] + k_spin[p6b] + k_spin[h1b] + k_spin[h2b] + k_spin[h3b] != 12))
      continue;
    if(!(k_spin[p4b] + k_spin[p5b] + k_spin[p6b] == k_spin[h1b] + k_spin[h2b] + k_spin[h3b]))
      continue;
    if(!(k_spin[p4b] == k_spin[h1b])) continue;
    if(!(k_range[p4b] > 0 && k_range[p5b] > 0 && k_range[p6b] > 0 && k_range[h1b] > 0 &&
         k_range[h2b] > 0 && k_range[h3b] > 0))
      continue;
    ia6_enabled[ia6] = true;
  }

  int idx_offset = 0;
  for(int ia6 = 0; ia6 < 9; ia6++) {
    if(!ia6_enabled[ia6]) continue;
    auto [p4b, p5b, p6b, h1b, h2b, h3b] = a3_s1[ia6];
    auto ref = std::make_tuple(t_p4b, t_p5b, t_p6b, t_h1b, t_h2b, t_h3b);
    if(ref == std::make_tuple(p4b, p5b, p6b, h1b, h2b, h3b)) df_simple_s1_exec[0] = idx_offset;
    if(ref == std::make_tuple(p4b, p5b, p6b, h2b, h1b, h3b)) df_simple_s1_exec[1] = idx_offset;
    if(ref == std::make_tuple(p4b, p5b, p6b, h2b, h3b, h1b)) df_simple_s1_exec[2] = idx_offset;
    if(ref == std::make_tuple(p5b, p4b, p6b, h1b, h2b, h3b)) df_simple_s1_exec[3] = idx_offset;
    if(ref == std::make_tuple(p5b, p4b, p6b, h2b, h1b, h3b)) df_simple_s1_exec[4] = idx_offset;
    if(ref == std::make_tuple(p5b, p4b, p6b, h2b, h3b, h1b)) df_simple_s1_exec[5] = idx_offset;
    if(ref == std::make_tuple(p5b, p6b, p4b, h1b, h2b, h3b)) df_simple_s1_exec[6] = idx_offset;
    if(ref == std::make_tuple(p5b, p6b, p4b, h2b, h1b, h3b)) df_simple_s1_exec[7] = idx_offset;
    if(ref == std::make_tuple(p5b, p6b, p4b, h2b, h3b, h1b)) df_simple_s1_exec[8] = idx_offset;
    if(comm_elems)
      *comm_elems += k_range[p4b] * k_range[h1b] +
                     k_range[p5b] * k_range[p6b] * k_range[h2b] * k_range[h3b];
    idx_offset++;
  }
  *df_num_s1_enabled = idx_offset;
  if(!data) return;

  // t1: d_t1.get({p4b-noab, h1b}) then perm {1,0}  (singles.hpp:182-233)
  std::vector<double> k_a, k_b;
  idx_offset = 0;
  for(int ia6 = 0; ia6 < 9; ia6++) {
    if(!ia6_enabled[ia6]) continue;
    auto [p4b, p5b, p6b, h1b, h2b, h3b] = a3_s1[ia6];
    (void) p5b; (void) p6b; (void) h2b; (void) h3b;
    const bool  vt[2] = {true, false};
    const Index bt[2] = {p4b, h1b};
    block_get(P, K_T1, P.t1, vt, bt, 2, k_a);
    int perm[2] = {1, 0};
    int size[2] = {(int) k_range[p4b], (int) k_range[h1b]};
    hptt_permute(perm, 2, 1.0, k_a.data(), size, df_T_s1_t1 + idx_offset * s1_max_dima);
    idx_offset++;
  }
  // v2: v2ijab.get({h3b,h2b,p6b-noab,p5b-noab}) then perm {3,2,1,0}  (singles.hpp:288-336)
  idx_offset = 0;
  for(int ia6 = 0; ia6 < 9; ia6++) {
    if(!ia6_enabled[ia6]) continue;
    auto [p4b, p5b, p6b, h1b, h2b, h3b] = a3_s1[ia6];
    (void) p4b; (void) h1b;
    const bool  vt[4] = {false, false, true, true};
    const Index bt[4] = {h3b, h2b, p6b, p5b};
    block_get(P, K_IJAB, P.v2ijab, vt, bt, 4, k_b);
    int perm[4] = {3, 2, 1, 0};
    int size[4] = {(int) k_range[h3b], (int) k_range[h2b], (int) k_range[p6b],
                   (int) k_range[p5b]};
    hptt_permute(perm, 4, 1.0, k_b.data(), size, df_T_s1_v2 + idx_offset * s1_max_dimb);
    idx_offset++;
  }
  (void) noab;
}

// --------------------------------------------------------------------------
// d1 gather: ccsd_t_all_fused_doubles1.hpp:37-307
// --------------------------------------------------------------------------
void data_d1_new(bool is_restricted, const SOProblem& P, size_t t_h1b, size_t t_h2b, size_t t_h3b,
                 size_t t_p4b, size_t t_p5b, size_t t_p6b, size_t max_d1_kernels_pertask,
                 size_t size_T_d1_t2, size_t size_T_d1_v2, double* df_T_d1_t2, double* df_T_d1_v2,
                 int* host_d1_size_h7b, int* df_simple_d1_size, int* df_simple_d1_exec,
                 int* df_num_d1_enabled, bool data, size_t* comm_elems) {
  const Index noab = P.noab;
  const auto& k_spin = P.k_spin;
  const auto& k_range = P.k_range;
  Tup6 a3_d1[] = {std::make_tuple(t_p4b, t_p5b, t_p6b, t_h1b, t_h2b, t_h3b),
                  std::make_tuple(t_p4b, t_p5b, t_p6b, t_h2b, t_h1b, t_h3b),
                  std::make_tuple(t_p4b, t_p5b, t_p6b, t_h3b, t_h1b, t_h2b),
                  std::make_tuple(t_p5b, t_p6b, t_p4b, t_h1b, t_h2b, t_h3b),
                  std::make_tuple(t_p5b, t_p6b, t_p4b, t_h2b, t_h1b, t_h3b),
                  std::make_tuple(t_p5b, t_p6b, t_p4b, t_h3b, t_h1b, t_h2b),
                  std::make_tuple(t_p4b, t_p6b, t_p5b, t_h1b, t_h2b, t_h3b),
                  std::make_tuple(t_p4b, t_p6b, t_p5b, t_h2b, t_h1b, t_h3b),
                  std::make_tuple(t_p4b, t_p6b, t_p5b, t_h3b, t_h1b, t_h2b)};
  for(int ia6 = 0; ia6 < 8; ia6++) {
    if(std::get<0>(a3_d1[ia6]) != 0) {
      for(int ja6 = ia6 + 1; ja6 < 9; ja6++)
        if(a3_d1[ia6] == a3_d1[ja6]) a3_d1[ja6] = zero6();
    }
  }
  const size_t max_dima = size_T_d1_t2 / max_d1_kernels_pertask;
  const size_t max_dimb = size_T_d1_v2 / max_d1_kernels_pertask;

  for(Index h7b = 0; h7b < noab; h7b++) {
    df_simple_d1_size[0 + h7b * 7] = (int) k_range[t_h1b];
    df_simple_d1_size[1 + h7b * 7] = (int) k_range[t_h2b];
    df_simple_d1_size[2 + h7b * 7] = (int) k_range[t_h3b];
    df_simple_d1_size[3 + h7b * 7] = (int) k_range[h7b];
    df_simple_d1_size[4 + h7b * 7] = (int) k_range[t_p4b];
    df_simple_d1_size[5 + h7b * 7] = (int) k_range[t_p5b];
    df_simple_d1_size[6 + h7b * 7] = (int) k_range[t_p6b];
    if(host_d1_size_h7b) host_d1_size_h7b[h7b] = (int) k_range[h7b];
  }

  std::vector<char> ia6_enabled(9 * noab, 0);
  for(int ia6 = 0; ia6 < 9; ia6++) {
    auto [p4b, p5b, p6b, h1b, h2b, h3b] = a3_d1[ia6];
    if(!((p4b <= p5b) && (h2b <= h3b) && p4b != 0)) continue;
    if(is_restricted &&
       !(k_spin[p4b] + k_spin[p5b] + k_spin[p6b] + k_spin[h1b] + k_spin[h2b] + k_spin[h3b] != 12))
      continue;
    if(!(k_spin[p4b] + k_spin[p5b] + k_spin[p6b] == k_spin[h1b] + k_spin[h2b] + k_spin[h3b]))
      continue;
    for(Index h7b = 0; h7b < noab; h7b++) {
      if(!(k_spin[p4b] + k_spin[p5b] == k_spin[h1b] + k_spin[h7b])) continue;
      if(!(k_range[p4b] > 0 && k_range[p5b] > 0 && k_range[p6b] > 0 && k_range[h1b] > 0 &&
           k_range[h2b] > 0 && k_range[h3b] > 0))
        continue;
      if(!(h7b <= p6b)) continue;
      ia6_enabled[ia6 * noab + h7b] = 1;
    }
  }

  int idx_offset = 0;
  for(int ia6 = 0; ia6 < 9; ia6++) {
    auto [p4b, p5b, p6b, h1b, h2b, h3b] = a3_d1[ia6];
    auto ref = std::make_tuple(t_p4b, t_p5b, t_p6b, t_h1b, t_h2b, t_h3b);
    for(Index h7b = 0; h7b < noab; h7b++) {
      if(!ia6_enabled[ia6 * noab + h7b]) continue;
      if(ref == std::make_tuple(p4b, p5b, p6b, h1b, h2b, h3b)) df_simple_d1_exec[0 + h7b * 9] = idx_offset;
      if(ref == std::make_tuple(p4b, p5b, p6b, h2b, h1b, h3b)) df_simple_d1_exec[1 + h7b * 9] = idx_offset;
      if(ref == std::make_tuple(p4b, p5b, p6b, h2b, h3b, h1b)) df_simple_d1_exec[2 + h7b * 9] = idx_offset;
      if(ref == std::make_tuple(p6b, p4b, p5b, h1b, h2b, h3b)) df_simple_d1_exec[3 + h7b * 9] = idx_offset;
      if(ref == std::make_tuple(p6b, p4b, p5b, h2b, h1b, h3b)) df_simple_d1_exec[4 + h7b * 9] = idx_offset;
      if(ref == std::make_tuple(p6b, p4b, p5b, h2b, h3b, h1b)) df_simple_d1_exec[5 + h7b * 9] = idx_offset;
      if(ref == std::make_tuple(p4b, p6b, p5b, h1b, h2b, h3b)) df_simple_d1_exec[6 + h7b * 9] = idx_offset;
      if(ref == std::make_tuple(p4b, p6b, p5b, h2b, h1b, h3b)) df_simple_d1_exec[7 + h7b * 9] = idx_offset;
      if(ref == std::make_tuple(p4b, p6b, p5b, h2b, h3b, h1b)) df_simple_d1_exec[8 + h7b * 9] = idx_offset;
      if(comm_elems)
        *comm_elems += k_range[h7b] * (k_range[p4b] * k_range[p5b] * k_range[h1b] +
                                       k_range[p6b] * k_range[h2b] * k_range[h3b]);
      idx_offset++;
    }
  }
  *df_num_d1_enabled = idx_offset;
  if(!data) return;

  std::vector<double> k_a, k_a_sort, k_b;
  idx_offset = 0;
  for(int ia6 = 0; ia6 < 9; ia6++) {
    auto [p4b, p5b, p6b, h1b, h2b, h3b] = a3_d1[ia6];
    (void) p6b; (void) h2b; (void) h3b;
    for(Index h7b = 0; h7b < noab; h7b++) {
      if(!ia6_enabled[ia6 * noab + h7b]) continue;
      size_t dima = k_range[h7b] * k_range[p4b] * k_range[p5b] * k_range[h1b];
      k_a_sort.assign(dima, 0.0);
      const bool vt[4] = {true, true, false, false};
      if(h7b < h1b) {  // doubles1.hpp:218-232
        const Index bt[4] = {p4b, p5b, h7b, h1b};
        block_get(P, K_T2, P.t2, vt, bt, 4, k_a);
        int perm[4] = {3, 1, 0, 2};
        int size[4] = {(int) k_range[p4b], (int) k_range[p5b], (int) k_range[h7b],
                       (int) k_range[h1b]};
        hptt_permute(perm, 4, -1.0, k_a.data(), size, k_a_sort.data());
      }
      if(h1b <= h7b) {  // doubles1.hpp:233-247
        const Index bt[4] = {p4b, p5b, h1b, h7b};
        block_get(P, K_T2, P.t2, vt, bt, 4, k_a);
        int perm[4] = {2, 1, 0, 3};
        int size[4] = {(int) k_range[p4b], (int) k_range[p5b], (int) k_range[h1b],
                       (int) k_range[h7b]};
        hptt_permute(perm, 4, 1.0, k_a.data(), size, k_a_sort.data());
      }
      std::copy(k_a_sort.begin(), k_a_sort.end(), df_T_d1_t2 + idx_offset * max_dima);
      idx_offset++;
    }
  }
  idx_offset = 0;
  for(int ia6 = 0; ia6 < 9; ia6++) {
    auto [p4b, p5b, p6b, h1b, h2b, h3b] = a3_d1[ia6];
    (void) p4b; (void) p5b; (void) h1b;
    for(Index h7b = 0; h7b < noab; h7b++) {
      if(!ia6_enabled[ia6 * noab + h7b]) continue;
      // v2ijka.get({h2b,h3b,h7b,p6b-noab}) perm {2,3,0,1}  (doubles1.hpp:278-290)
      const bool  vt[4] = {false, false, false, true};
      const Index bt[4] = {h2b, h3b, h7b, p6b};
      block_get(P, K_IJKA, P.v2ijka, vt, bt, 4, k_b);
      int perm[4] = {2, 3, 0, 1};
      int size[4] = {(int) k_range[h2b], (int) k_range[h3b], (int) k_range[h7b],
                     (int) k_range[p6b]};
      hptt_permute(perm, 4, 1.0, k_b.data(), size, df_T_d1_v2 + idx_offset * max_dimb);
      idx_offset++;
    }
  }
}

// --------------------------------------------------------------------------
// d2 gather: ccsd_t_all_fused_doubles2.hpp:37-360
// --------------------------------------------------------------------------
void data_d2_new(bool is_restricted, const SOProblem& P, size_t t_h1b, size_t t_h2b, size_t t_h3b,
                 size_t t_p4b, size_t t_p5b, size_t t_p6b, size_t max_d2_kernels_pertask,
                 size_t size_T_d2_t2, size_t size_T_d2_v2, double* df_T_d2_t2, double* df_T_d2_v2,
                 int* host_d2_size_p7b, int* df_simple_d2_size, int* df_simple_d2_exec,
                 int* df_num_d2_enabled, bool data, size_t* comm_elems) {
  const Index noab = P.noab;
  const Index nvab = P.nvab;
  const auto& k_spin = P.k_spin;
  const auto& k_range = P.k_range;
  Tup6 a3_d2[] = {std::make_tuple(t_p4b, t_p5b, t_p6b, t_h1b, t_h2b, t_h3b),
                  std::make_tuple(t_p4b, t_p5b, t_p6b, t_h2b, t_h3b, t_h1b),
                  std::make_tuple(t_p4b, t_p5b, t_p6b, t_h1b, t_h3b, t_h2b),
                  std::make_tuple(t_p5b, t_p4b, t_p6b, t_h1b, t_h2b, t_h3b),
                  std::make_tuple(t_p5b, t_p4b, t_p6b, t_h2b, t_h3b, t_h1b),
                  std::make_tuple(t_p5b, t_p4b, t_p6b, t_h1b, t_h3b, t_h2b),
                  std::make_tuple(t_p6b, t_p4b, t_p5b, t_h1b, t_h2b, t_h3b),
                  std::make_tuple(t_p6b, t_p4b, t_p5b, t_h2b, t_h3b, t_h1b),
                  std::make_tuple(t_p6b, t_p4b, t_p5b, t_h1b, t_h3b, t_h2b)};
  for(int ia6 = 0; ia6 < 8; ia6++) {
    if(std::get<0>(a3_d2[ia6]) != 0) {
      for(int ja6 = ia6 + 1; ja6 < 9; ja6++)
        if(a3_d2[ia6] == a3_d2[ja6]) a3_d2[ja6] = zero6();
    }
  }
  const size_t max_dima2 = size_T_d2_t2 / max_d2_kernels_pertask;
  const size_t max_dimb2 = size_T_d2_v2 / max_d2_kernels_pertask;

  for(Index p7b = noab; p7b < noab + nvab; p7b++) {
    df_simple_d2_size[0 + (p7b - noab) * 7] = (int) k_range[t_h1b];
    df_simple_d2_size[1 + (p7b - noab) * 7] = (int) k_range[t_h2b];
    df_simple_d2_size[2 + (p7b - noab) * 7] = (int) k_range[t_h3b];
    df_simple_d2_size[3 + (p7b - noab) * 7] = (int) k_range[t_p4b];
    df_simple_d2_size[4 + (p7b - noab) * 7] = (int) k_range[t_p5b];
    df_simple_d2_size[5 + (p7b - noab) * 7] = (int) k_range[t_p6b];
    df_simple_d2_size[6 + (p7b - noab) * 7] = (int) k_range[p7b];
    if(host_d2_size_p7b) host_d2_size_p7b[p7b - noab] = (int) k_range[p7b];
  }

  std::vector<char> ia6_enabled(9 * nvab, 0);
  for(int ia6 = 0; ia6 < 9; ia6++) {
    auto [p4b, p5b, p6b, h1b, h2b, h3b] = a3_d2[ia6];
    if(!((p5b <= p6b) && (h1b <= h2b) && p4b != 0)) continue;
    if(is_restricted &&
       !(k_spin[p4b] + k_spin[p5b] + k_spin[p6b] + k_spin[h1b] + k_spin[h2b] + k_spin[h3b] != 12))
      continue;
    if(!(k_spin[p4b] + k_spin[p5b] + k_spin[p6b] == k_spin[h1b] + k_spin[h2b] + k_spin[h3b]))
      continue;
    for(Index p7b = noab; p7b < noab + nvab; p7b++) {
      if(!(k_spin[p4b] + k_spin[p7b] == k_spin[h1b] + k_spin[h2b])) continue;
      if(!(k_range[p4b] > 0 && k_range[p5b] > 0 && k_range[p6b] > 0 && k_range[h1b] > 0 &&
           k_range[h2b] > 0 && k_range[h3b] > 0))
        continue;
      if(!(h3b <= p7b)) continue;
      ia6_enabled[ia6 * nvab + p7b - noab] = 1;
    }
  }

  int idx_offset = 0;
  for(int ia6 = 0; ia6 < 9; ia6++) {
    auto [p4b, p5b, p6b, h1b, h2b, h3b] = a3_d2[ia6];
    auto ref = std::make_tuple(t_p4b, t_p5b, t_p6b, t_h1b, t_h2b, t_h3b);
    for(Index p7b = noab; p7b < noab + nvab; p7b++) {
      if(!ia6_enabled[ia6 * nvab + p7b - noab]) continue;
      int* ex = df_simple_d2_exec + (p7b - noab) * 9;
      if(ref == std::make_tuple(p4b, p5b, p6b, h1b, h2b, h3b)) ex[0] = idx_offset;
      if(ref == std::make_tuple(p4b, p5b, p6b, h3b, h1b, h2b)) ex[1] = idx_offset;
      if(ref == std::make_tuple(p4b, p5b, p6b, h1b, h3b, h2b)) ex[2] = idx_offset;
      if(ref == std::make_tuple(p5b, p4b, p6b, h1b, h2b, h3b)) ex[3] = idx_offset;
      if(ref == std::make_tuple(p5b, p4b, p6b, h3b, h1b, h2b)) ex[4] = idx_offset;
      if(ref == std::make_tuple(p5b, p4b, p6b, h1b, h3b, h2b)) ex[5] = idx_offset;
      if(ref == std::make_tuple(p5b, p6b, p4b, h1b, h2b, h3b)) ex[6] = idx_offset;
      if(ref == std::make_tuple(p5b, p6b, p4b, h3b, h1b, h2b)) ex[7] = idx_offset;
      if(ref == std::make_tuple(p5b, p6b, p4b, h1b, h3b, h2b)) ex[8] = idx_offset;
      if(comm_elems)
        *comm_elems += k_range[p7b] * (k_range[p4b] * k_range[h1b] * k_range[h2b] +
                                       k_range[p5b] * k_range[p6b] * k_range[h3b]);
      idx_offset++;
    }
  }
  *df_num_d2_enabled = idx_offset;
  if(!data) return;

  std::vector<double> k_a, k_a_sort, k_b;
  idx_offset = 0;
  for(int ia6 = 0; ia6 < 9; ia6++) {
    auto [p4b, p5b, p6b, h1b, h2b, h3b] = a3_d2[ia6];
    (void) p5b; (void) p6b; (void) h3b;
    for(Index p7b = noab; p7b < noab + nvab; p7b++) {
      if(!ia6_enabled[ia6 * nvab + p7b - noab]) continue;
      size_t dima = k_range[p7b] * k_range[p4b] * k_range[h1b] * k_range[h2b];
      k_a_sort.assign(dima, 0.0);
      const bool vt[4] = {true, true, false, false};
      if(p7b < p4b) {  // doubles2.hpp:211-225
        const Index bt[4] = {p7b, p4b, h1b, h2b};
        block_get(P, K_T2, P.t2, vt, bt, 4, k_a);
        int perm[4] = {3, 2, 1, 0};
        int size[4] = {(int) k_range[p7b], (int) k_range[p4b], (int) k_range[h1b],
                       (int) k_range[h2b]};
        hptt_permute(perm, 4, -1.0, k_a.data(), size, k_a_sort.data());
      }
      if(p4b <= p7b) {  // doubles2.hpp:226-239
        const Index bt[4] = {p4b, p7b, h1b, h2b};
        block_get(P, K_T2, P.t2, vt, bt, 4, k_a);
        int perm[4] = {3, 2, 0, 1};
        int size[4] = {(int) k_range[p4b], (int) k_range[p7b], (int) k_range[h1b],
                       (int) k_range[h2b]};
        hptt_permute(perm, 4, 1.0, k_a.data(), size, k_a_sort.data());
      }
      std::copy(k_a_sort.begin(), k_a_sort.end(), df_T_d2_t2 + idx_offset * max_dima2);
      idx_offset++;
    }
  }
  idx_offset = 0;
  for(int ia6 = 0; ia6 < 9; ia6++) {
    auto [p4b, p5b, p6b, h1b, h2b, h3b] = a3_d2[ia6];
    (void) p4b; (void) h1b; (void) h2b;
    for(Index p7b = noab; p7b < noab + nvab; p7b++) {
      if(!ia6_enabled[ia6 * nvab + p7b - noab]) continue;
      // v2iabc.get({h3b,p7b-noab,p5b-noab,p6b-noab}) perm {2,3,0,1}  (doubles2.hpp:331-343)
      const bool  vt[4] = {false, true, true, true};
      const Index bt[4] = {h3b, p7b, p5b, p6b};
      block_get(P, K_IABC, P.v2iabc, vt, bt, 4, k_b);
      int perm[4] = {2, 3, 0, 1};
      int size[4] = {(int) k_range[h3b], (int) k_range[p7b], (int) k_range[p5b],
                     (int) k_range[p6b]};
      hptt_permute(perm, 4, 1.0, k_b.data(), size, df_T_d2_v2 + idx_offset * max_dimb2);
      idx_offset++;
    }
  }
}

// Buffer sizing of the fused driver, ccsd_t_fused_driver.hpp:176-195 (tilesize_opt = true).
struct Sizes {
  size_t max_d1_kernels_pertask, max_d2_kernels_pertask;
  size_t s1_t1, s1_v2, d1_t2, d1_v2, d2_t2, d2_v2;
};

Sizes driver_sizes(const SOProblem& P) {
  size_t max_pdim = 0, max_hdim = 0;
  for(Index b = P.noab; b < P.noab + P.nvab; b++) max_pdim = std::max(max_pdim, P.k_range[b]);
  for(Index b = 0; b < P.noab; b++) max_hdim = std::max(max_hdim, P.k_range[b]);
  Sizes s;
  size_t noa = 0, nva = 0;
  for(Index b = 0; b < P.noab; b++) noa += (P.k_spin[b] == 1);
  for(Index b = P.noab; b < P.noab + P.nvab; b++) nva += (P.k_spin[b] == 1);
  s.max_d1_kernels_pertask = 9 * noa;
  s.max_d2_kernels_pertask = 9 * nva;
  s.s1_t1 = 9 * max_pdim * max_hdim;
  s.s1_v2 = 9 * (max_pdim * max_pdim) * (max_hdim * max_hdim);
  s.d1_t2 = s.max_d1_kernels_pertask * (max_pdim * max_pdim) * (max_hdim * max_hdim);
  s.d1_v2 = s.max_d1_kernels_pertask * max_pdim * (max_hdim * max_hdim * max_hdim);
  s.d2_t2 = s.max_d2_kernels_pertask * (max_pdim * max_pdim) * (max_hdim * max_hdim);
  s.d2_v2 = s.max_d2_kernels_pertask * (max_pdim * max_pdim * max_pdim) * max_hdim;
  return s;
}

// Packed buffers + metadata for one task (pinned host buffers of the reference).
struct TaskBuffers {
  std::vector<double> s1_t1, s1_v2, d1_t2, d1_v2, d2_t2, d2_v2;
  std::vector<int>    s1_size, s1_exec, d1_size, d1_exec, d2_size, d2_exec;
  int                 n_s1 = 0, n_d1 = 0, n_d2 = 0;
};

void gather_task(bool is_restricted, const SOProblem& P, const Sizes& S, const size_t* t,
                 TaskBuffers& B) {
  B.s1_t1.assign(S.s1_t1, 0.0);
  B.s1_v2.assign(S.s1_v2, 0.0);
  B.d1_t2.assign(S.d1_t2, 0.0);
  B.d1_v2.assign(S.d1_v2, 0.0);
  B.d2_t2.assign(S.d2_t2, 0.0);
  B.d2_v2.assign(S.d2_v2, 0.0);
  B.s1_size.assign(6, 0);
  B.d1_size.assign(7 * P.noab, 0);
  B.d2_size.assign(7 * P.nvab, 0);
  // resets, ccsd_t_all_fused_cpu.hpp:71-73
  B.s1_exec.assign(9, -1);
  B.d1_exec.assign(9 * P.noab, -1);
  B.d2_exec.assign(9 * P.nvab, -1);
  data_s1_new(is_restricted, P, t[0], t[1], t[2], t[3], t[4], t[5], S.s1_t1, S.s1_v2,
              B.s1_size.data(), B.s1_exec.data(), B.s1_t1.data(), B.s1_v2.data(), &B.n_s1, true,
              nullptr);
  data_d1_new(is_restricted, P, t[0], t[1], t[2], t[3], t[4], t[5], S.max_d1_kernels_pertask,
              S.d1_t2, S.d1_v2, B.d1_t2.data(), B.d1_v2.data(), nullptr, B.d1_size.data(),
              B.d1_exec.data(), &B.n_d1, true, nullptr);
  data_d2_new(is_restricted, P, t[0], t[1], t[2], t[3], t[4], t[5], S.max_d2_kernels_pertask,
              S.d2_t2, S.d2_v2, B.d2_t2.data(), B.d2_v2.data(), nullptr, B.d2_size.data(),
              B.d2_exec.data(), &B.n_d2, true, nullptr);
}

// --------------------------------------------------------------------------
// Per-task CPU kernel: ccsd_t_all_fused_cpu.hpp:102-623
// Returns the two values the reference adds to energy_l (factor included).
// --------------------------------------------------------------------------
void task_kernel_cpu(const SOProblem& P, const Sizes& S, const size_t* t, double factor,
                     const TaskBuffers& B, double* out_e) {
  const size_t t_h1b = t[0], t_h2b = t[1], t_h3b = t[2], t_p4b = t[3], t_p5b = t[4], t_p6b = t[5];
  const auto&  k_range = P.k_range;
  size_t base_size_h1b = k_range[t_h1b], base_size_h2b = k_range[t_h2b],
         base_size_h3b = k_range[t_h3b], base_size_p4b = k_range[t_p4b],
         base_size_p5b = k_range[t_p5b], base_size_p6b = k_range[t_p6b];
  const size_t max_dim_s1_t1 = S.s1_t1 / 9;
  const size_t max_dim_s1_v2 = S.s1_v2 / 9;
  const size_t max_dim_d1_t2 = S.d1_t2 / S.max_d1_kernels_pertask;
  const size_t max_dim_d1_v2 = S.d1_v2 / S.max_d1_kernels_pertask;
  const size_t max_dim_d2_t2 = S.d2_t2 / S.max_d2_kernels_pertask;
  const size_t max_dim_d2_v2 = S.d2_v2 / S.max_d2_kernels_pertask;

  const double* host_evl_sorted_h1b = &P.k_evl_sorted[P.k_offset[t_h1b]];
  const double* host_evl_sorted_h2b = &P.k_evl_sorted[P.k_offset[t_h2b]];
  const double* host_evl_sorted_h3b = &P.k_evl_sorted[P.k_offset[t_h3b]];
  const double* host_evl_sorted_p4b = &P.k_evl_sorted[P.k_offset[t_p4b]];
  const double* host_evl_sorted_p5b = &P.k_evl_sorted[P.k_offset[t_p5b]];
  const double* host_evl_sorted_p6b = &P.k_evl_sorted[P.k_offset[t_p6b]];

  size_t size_tensor_t3 =
    base_size_h3b * base_size_h2b * base_size_h1b * base_size_p6b * base_size_p5b * base_size_p4b;
  std::vector<double> host_t3_d(size_tensor_t3, 0.0), host_t3_s(size_tensor_t3, 0.0);
  double* t3d = host_t3_d.data();
  double* t3s = host_t3_s.data();

  // d1, ccsd_t_all_fused_cpu.hpp:118-273
  for(size_t idx_noab = 0; idx_noab < P.noab; idx_noab++) {
    const int* fl = &B.d1_exec[idx_noab * 9];
    const int* sz = &B.d1_size[idx_noab * 7];
    const int  H1 = sz[0], H2 = sz[1], H3 = sz[2], H7 = sz[3], P4 = sz[4], P5 = sz[5], P6 = sz[6];
    const double* t2_[9];
    const double* v2_[9];
    for(int e = 0; e < 9; e++) {
      t2_[e] = B.d1_t2.data() + max_dim_d1_t2 * (std::ptrdiff_t) fl[e];
      v2_[e] = B.d1_v2.data() + max_dim_d1_v2 * (std::ptrdiff_t) fl[e];
    }
#ifdef _OPENMP
#pragma omp parallel for collapse(6)
#endif
    for(int h3 = 0; h3 < H3; h3++)
      for(int h2 = 0; h2 < H2; h2++)
        for(int h1 = 0; h1 < H1; h1++)
          for(int p6 = 0; p6 < P6; p6++)
            for(int p5 = 0; p5 < P5; p5++)
              for(int p4 = 0; p4 < P4; p4++) {
                int idx = h3 + (h2 + (h1 + (p6 + (p5 + p4 * P5) * P6) * H1) * H2) * H3;
                for(int h7 = 0; h7 < H7; h7++) {
                  if(fl[0] >= 0) t3d[idx] -= t2_[0][h7 + (p4 + (p5 + h1 * P5) * P4) * H7] * v2_[0][h3 + (h2 + (p6 + h7 * P6) * H2) * H3];
                  if(fl[1] >= 0) t3d[idx] += t2_[1][h7 + (p4 + (p5 + h2 * P5) * P4) * H7] * v2_[1][h3 + (h1 + (p6 + h7 * P6) * H1) * H3];
                  if(fl[2] >= 0) t3d[idx] -= t2_[2][h7 + (p4 + (p5 + h3 * P5) * P4) * H7] * v2_[2][h2 + (h1 + (p6 + h7 * P6) * H1) * H2];
                  if(fl[3] >= 0) t3d[idx] -= t2_[3][h7 + (p5 + (p6 + h1 * P6) * P5) * H7] * v2_[3][h3 + (h2 + (p4 + h7 * P4) * H2) * H3];
                  if(fl[4] >= 0) t3d[idx] += t2_[4][h7 + (p5 + (p6 + h2 * P6) * P5) * H7] * v2_[4][h3 + (h1 + (p4 + h7 * P4) * H1) * H3];
                  if(fl[5] >= 0) t3d[idx] -= t2_[5][h7 + (p5 + (p6 + h3 * P6) * P5) * H7] * v2_[5][h2 + (h1 + (p4 + h7 * P4) * H1) * H2];
                  if(fl[6] >= 0) t3d[idx] += t2_[6][h7 + (p4 + (p6 + h1 * P6) * P4) * H7] * v2_[6][h3 + (h2 + (p5 + h7 * P5) * H2) * H3];
                  if(fl[7] >= 0) t3d[idx] -= t2_[7][h7 + (p4 + (p6 + h2 * P6) * P4) * H7] * v2_[7][h3 + (h1 + (p5 + h7 * P5) * H1) * H3];
                  if(fl[8] >= 0) t3d[idx] += t2_[8][h7 + (p4 + (p6 + h3 * P6) * P4) * H7] * v2_[8][h2 + (h1 + (p5 + h7 * P5) * H1) * H2];
                }
              }
  }

  // d2, ccsd_t_all_fused_cpu.hpp:276-431
  for(size_t idx_nvab = 0; idx_nvab < P.nvab; idx_nvab++) {
    const int* fl = &B.d2_exec[idx_nvab * 9];
    const int* sz = &B.d2_size[idx_nvab * 7];
    const int  H1 = sz[0], H2 = sz[1], H3 = sz[2], P4 = sz[3], P5 = sz[4], P6 = sz[5], P7 = sz[6];
    const double* t2_[9];
    const double* v2_[9];
    for(int e = 0; e < 9; e++) {
      t2_[e] = B.d2_t2.data() + max_dim_d2_t2 * (std::ptrdiff_t) fl[e];
      v2_[e] = B.d2_v2.data() + max_dim_d2_v2 * (std::ptrdiff_t) fl[e];
    }
#ifdef _OPENMP
#pragma omp parallel for collapse(6)
#endif
    for(int h3 = 0; h3 < H3; h3++)
      for(int h2 = 0; h2 < H2; h2++)
        for(int h1 = 0; h1 < H1; h1++)
          for(int p6 = 0; p6 < P6; p6++)
            for(int p5 = 0; p5 < P5; p5++)
              for(int p4 = 0; p4 < P4; p4++) {
                int idx = h3 + (h2 + (h1 + (p6 + (p5 + p4 * P5) * P6) * H1) * H2) * H3;
                for(int p7 = 0; p7 < P7; p7++) {
                  if(fl[0] >= 0) t3d[idx] -= t2_[0][p7 + (p4 + (h1 + h2 * H1) * P4) * P7] * v2_[0][p7 + (h3 + (p6 + p5 * P6) * H3) * P7];
                  if(fl[1] >= 0) t3d[idx] -= t2_[1][p7 + (p4 + (h2 + h3 * H2) * P4) * P7] * v2_[1][p7 + (h1 + (p6 + p5 * P6) * H1) * P7];
                  if(fl[2] >= 0) t3d[idx] += t2_[2][p7 + (p4 + (h1 + h3 * H1) * P4) * P7] * v2_[2][p7 + (h2 + (p6 + p5 * P6) * H2) * P7];
                  if(fl[3] >= 0) t3d[idx] += t2_[3][p7 + (p5 + (h1 + h2 * H1) * P5) * P7] * v2_[3][p7 + (h3 + (p6 + p4 * P6) * H3) * P7];
                  if(fl[4] >= 0) t3d[idx] += t2_[4][p7 + (p5 + (h2 + h3 * H2) * P5) * P7] * v2_[4][p7 + (h1 + (p6 + p4 * P6) * H1) * P7];
                  if(fl[5] >= 0) t3d[idx] -= t2_[5][p7 + (p5 + (h1 + h3 * H1) * P5) * P7] * v2_[5][p7 + (h2 + (p6 + p4 * P6) * H2) * P7];
                  if(fl[6] >= 0) t3d[idx] -= t2_[6][p7 + (p6 + (h1 + h2 * H1) * P6) * P7] * v2_[6][p7 + (h3 + (p5 + p4 * P5) * H3) * P7];
                  if(fl[7] >= 0) t3d[idx] -= t2_[7][p7 + (p6 + (h2 + h3 * H2) * P6) * P7] * v2_[7][p7 + (h1 + (p5 + p4 * P5) * H1) * P7];
                  if(fl[8] >= 0) t3d[idx] += t2_[8][p7 + (p6 + (h1 + h3 * H1) * P6) * P7] * v2_[8][p7 + (h2 + (p5 + p4 * P5) * H2) * P7];
                }
              }
  }

  // s1, ccsd_t_all_fused_cpu.hpp:434-580 (slots 7 and 9 add, as the code does)
  {
    const int* fl = B.s1_exec.data();
    const int* sz = B.s1_size.data();
    const int  H1 = sz[0], H2 = sz[1], H3 = sz[2], P4 = sz[3], P5 = sz[4], P6 = sz[5];
    const double* t1_[9];
    const double* v2_[9];
    for(int e = 0; e < 9; e++) {
      t1_[e] = B.s1_t1.data() + max_dim_s1_t1 * (std::ptrdiff_t) fl[e];
      v2_[e] = B.s1_v2.data() + max_dim_s1_v2 * (std::ptrdiff_t) fl[e];
    }
#ifdef _OPENMP
#pragma omp parallel for collapse(6)
#endif
    for(int h3 = 0; h3 < H3; h3++)
      for(int h2 = 0; h2 < H2; h2++)
        for(int h1 = 0; h1 < H1; h1++)
          for(int p6 = 0; p6 < P6; p6++)
            for(int p5 = 0; p5 < P5; p5++)
              for(int p4 = 0; p4 < P4; p4++) {
                int idx = h3 + (h2 + (h1 + (p6 + (p5 + p4 * P5) * P6) * H1) * H2) * H3;
                if(fl[0] >= 0) t3s[idx] += t1_[0][p4 + h1 * P4] * v2_[0][h3 + (h2 + (p6 + p5 * P6) * H2) * H3];
                if(fl[1] >= 0) t3s[idx] -= t1_[1][p4 + h2 * P4] * v2_[1][h3 + (h1 + (p6 + p5 * P6) * H1) * H3];
                if(fl[2] >= 0) t3s[idx] += t1_[2][p4 + h3 * P4] * v2_[2][h2 + (h1 + (p6 + p5 * P6) * H1) * H2];
                if(fl[3] >= 0) t3s[idx] -= t1_[3][p5 + h1 * P5] * v2_[3][h3 + (h2 + (p6 + p4 * P6) * H2) * H3];
                if(fl[4] >= 0) t3s[idx] += t1_[4][p5 + h2 * P5] * v2_[4][h3 + (h1 + (p6 + p4 * P6) * H1) * H3];
                if(fl[5] >= 0) t3s[idx] -= t1_[5][p5 + h3 * P5] * v2_[5][h2 + (h1 + (p6 + p4 * P6) * H1) * H2];
                if(fl[6] >= 0) t3s[idx] += t1_[6][p6 + h1 * P6] * v2_[6][h3 + (h2 + (p5 + p4 * P5) * H2) * H3];
                if(fl[7] >= 0) t3s[idx] -= t1_[7][p6 + h2 * P6] * v2_[7][h3 + (h1 + (p5 + p4 * P5) * H1) * H3];
                if(fl[8] >= 0) t3s[idx] += t1_[8][p6 + h3 * P6] * v2_[8][h2 + (h1 + (p5 + p4 * P5) * H1) * H2];
              }
  }

  // energies, ccsd_t_all_fused_cpu.hpp:586-623 (serial, same order).  The
  // reference carries these sums in double; here they are carried in 80-bit
  // long double so that the oracle's value is the exactly rounded sum of the
  // per-element terms (a serial double sum of ~1e8 terms drifts by ~1e-12
  // relative, more than the parity tolerance of an FP64 implementation).
  long double final_energy_1 = 0.0L, final_energy_2 = 0.0L;
  const int size_idx_h1 = (int) base_size_h1b, size_idx_h2 = (int) base_size_h2b,
            size_idx_h3 = (int) base_size_h3b, size_idx_p4 = (int) base_size_p4b,
            size_idx_p5 = (int) base_size_p5b, size_idx_p6 = (int) base_size_p6b;
  for(int idx_p4 = 0; idx_p4 < size_idx_p4; idx_p4++)
    for(int idx_p5 = 0; idx_p5 < size_idx_p5; idx_p5++)
      for(int idx_p6 = 0; idx_p6 < size_idx_p6; idx_p6++)
        for(int idx_h1 = 0; idx_h1 < size_idx_h1; idx_h1++)
          for(int idx_h2 = 0; idx_h2 < size_idx_h2; idx_h2++)
            for(int idx_h3 = 0; idx_h3 < size_idx_h3; idx_h3++) {
              int idx_t3 =
                idx_h3 + (idx_h2 + (idx_h1 + (idx_p6 + (idx_p5 + idx_p4 * size_idx_p5) * size_idx_p6) *
                                               size_idx_h1) * size_idx_h2) * size_idx_h3;
              double inner_factor = (host_evl_sorted_h3b[idx_h3] + host_evl_sorted_h2b[idx_h2] +
                                     host_evl_sorted_h1b[idx_h1] - host_evl_sorted_p6b[idx_p6] -
                                     host_evl_sorted_p5b[idx_p5] - host_evl_sorted_p4b[idx_p4]);
              final_energy_1 += factor * t3d[idx_t3] * (t3d[idx_t3]) / inner_factor;
              final_energy_2 += factor * t3d[idx_t3] * (t3d[idx_t3] + t3s[idx_t3]) / inner_factor;
            }
  out_e[0] = (double) final_energy_1;
  out_e[1] = (double) final_energy_2;
}

SOProblem make_problem(int noab, int nvab, const int* tiles, const int* spins,
                       const double* evl, const double* t1, const double* t2, const double* v2ijab,
                       const double* v2ijka, const double* v2iabc) {
  SOProblem P;
  P.noab = (Index) noab;
  P.nvab = (Index) nvab;
  size_t sum = 0;
  for(int b = 0; b < noab + nvab; b++) {
    P.k_range.push_back((size_t) tiles[b]);
    P.k_offset.push_back(sum);
    P.k_spin.push_back(spins[b]);
    sum += (size_t) tiles[b];
  }
  P.nO = 0;
  for(int b = 0; b < noab; b++) P.nO += (size_t) tiles[b];
  P.nV = sum - P.nO;
  if(evl) P.k_evl_sorted.assign(evl, evl + sum);
  P.t1 = t1;
  P.t2 = t2;
  P.v2ijab = v2ijab;
  P.v2ijka = v2ijka;
  P.v2iabc = v2iabc;
  return P;
}

// Task list of the fused driver, ccsd_t_fused_driver.hpp:368-395 (seq h3b variant).
void build_tasks(const SOProblem& P, bool is_restricted, std::vector<size_t>& tasks,
                 std::vector<double>& factors) {
  const Index noab = P.noab, nvab = P.nvab;
  const auto& k_spin = P.k_spin;
  for(size_t t_h1b = 0; t_h1b < noab; t_h1b++)
    for(size_t t_p4b = noab; t_p4b < noab + nvab; t_p4b++)
      for(size_t t_h2b = t_h1b; t_h2b < noab; t_h2b++)
        for(size_t t_p5b = t_p4b; t_p5b < noab + nvab; t_p5b++)
          for(size_t t_p6b = t_p5b; t_p6b < noab + nvab; t_p6b++)
            for(size_t t_h3b = t_h2b; t_h3b < noab; t_h3b++) {
              if((k_spin[t_p4b] + k_spin[t_p5b] + k_spin[t_p6b]) !=
                 (k_spin[t_h1b] + k_spin[t_h2b] + k_spin[t_h3b]))
                continue;
              if(is_restricted && (k_spin[t_p4b] + k_spin[t_p5b] + k_spin[t_p6b] + k_spin[t_h1b] +
                                   k_spin[t_h2b] + k_spin[t_h3b]) > 8)
                continue;
              double factor = 1.0;
              if(is_restricted) factor = 2.0;
              if((t_p4b == t_p5b) && (t_p5b == t_p6b)) factor /= 6.0;
              else if((t_p4b == t_p5b) || (t_p5b == t_p6b)) factor /= 2.0;
              if((t_h1b == t_h2b) && (t_h2b == t_h3b)) factor /= 6.0;
              else if((t_h1b == t_h2b) || (t_h2b == t_h3b)) factor /= 2.0;
              size_t tt[6] = {t_h1b, t_h2b, t_h3b, t_p4b, t_p5b, t_p6b};
              tasks.insert(tasks.end(), tt, tt + 6);
              factors.push_back(factor);
            }
}

}  // namespace

extern "C" {

// Number of energy tasks (6-tuples passing the spin filter), in driver order.
long oracle_num_tasks(int noab, int nvab, const int* tiles, const int* spins, int is_restricted) {
  SOProblem P = make_problem(noab, nvab, tiles, spins, nullptr, nullptr, nullptr, nullptr, nullptr,
                             nullptr);
  std::vector<size_t> t;
  std::vector<double> f;
  build_tasks(P, is_restricted != 0, t, f);
  return (long) f.size();
}

// Task list: tasks_out[6*n] = (h1b,h2b,h3b,p4b,p5b,p6b), factors_out[n].
long oracle_task_list(int noab, int nvab, const int* tiles, const int* spins, int is_restricted,
                      long* tasks_out, double* factors_out) {
  SOProblem P = make_problem(noab, nvab, tiles, spins, nullptr, nullptr, nullptr, nullptr, nullptr,
                             nullptr);
  std::vector<size_t> t;
  std::vector<double> f;
  build_tasks(P, is_restricted != 0, t, f);
  for(size_t i = 0; i < t.size(); i++) tasks_out[i] = (long) t[i];
  for(size_t i = 0; i < f.size(); i++) factors_out[i] = f[i];
  return (long) f.size();
}

// Analytic flop count, CCSD_T_Fused_Driver::calculate_performance_ops
// (ccsd_t_fused_driver.hpp:548-638) -> fused_common.hpp:34-128, 131-261.
// out[0..2] = s1, d1, d2 totals; per_task (optional) = per-task total.
// comm_elems_out (optional) = per-task gathered element count (info_only).
long double oracle_flops(int noab, int nvab, const int* tiles, const int* spins, int is_restricted,
                         long double* out3, double* per_task, double* comm_elems_out) {
  SOProblem P = make_problem(noab, nvab, tiles, spins, nullptr, nullptr, nullptr, nullptr, nullptr,
                             nullptr);
  std::vector<size_t> tasks;
  std::vector<double> factors;
  build_tasks(P, is_restricted != 0, tasks, factors);
  std::vector<int> s1_size(6), s1_exec(9), d1_size(7 * noab), d1_exec(9 * noab),
    d2_size(7 * nvab), d2_exec(9 * nvab);
  long double tot_s1 = 0, tot_d1 = 0, tot_d2 = 0;
  for(size_t k = 0; k < factors.size(); k++) {
    const size_t* t = &tasks[6 * k];
    std::fill(s1_exec.begin(), s1_exec.end(), -1);
    std::fill(d1_exec.begin(), d1_exec.end(), -1);
    std::fill(d2_exec.begin(), d2_exec.end(), -1);
    int    n1 = 0, n2 = 0, n3 = 0;
    size_t comm = 0;
    data_s1_new(is_restricted != 0, P, t[0], t[1], t[2], t[3], t[4], t[5], 9, 9, s1_size.data(),
                s1_exec.data(), nullptr, nullptr, &n1, false, &comm);
    data_d1_new(is_restricted != 0, P, t[0], t[1], t[2], t[3], t[4], t[5], 1, 1, 1, nullptr,
                nullptr, nullptr, d1_size.data(), d1_exec.data(), &n2, false, &comm);
    data_d2_new(is_restricted != 0, P, t[0], t[1], t[2], t[3], t[4], t[5], 1, 1, 1, nullptr,
                nullptr, nullptr, d2_size.data(), d2_exec.data(), &n3, false, &comm);
    // helper_calculate_num_ops, fused_common.hpp:131-261
    long double ops_s1 = 0, ops_d1 = 0, ops_d2 = 0;
    long double base_s1 = (long double) s1_size[0] * (long double) s1_size[1] *
                          (long double) s1_size[2] * (long double) s1_size[3] *
                          (long double) s1_size[4] * (long double) s1_size[5] * 2;
    for(int e = 0; e < 9; e++)
      if(s1_exec[e] >= 0) ops_s1 += base_s1;
    for(int b = 0; b < noab; b++) {
      long double base = 2;
      for(int d = 0; d < 7; d++) base *= (long double) d1_size[d + b * 7];
      for(int e = 0; e < 9; e++)
        if(d1_exec[e + b * 9] >= 0) ops_d1 += base;
    }
    for(int b = 0; b < nvab; b++) {
      long double base = 2;
      for(int d = 0; d < 7; d++) base *= (long double) d2_size[d + b * 7];
      for(int e = 0; e < 9; e++)
        if(d2_exec[e + b * 9] >= 0) ops_d2 += base;
    }
    tot_s1 += ops_s1;
    tot_d1 += ops_d1;
    tot_d2 += ops_d2;
    if(per_task) per_task[k] = (double) (ops_s1 + ops_d1 + ops_d2);
    if(comm_elems_out) comm_elems_out[k] = (double) comm;
  }
  if(out3) {
    out3[0] = tot_s1;
    out3[1] = tot_d1;
    out3[2] = tot_d2;
  }
  return tot_s1 + tot_d1 + tot_d2;
}

// Exec flags of one task as the gathers produce them (for checking the
// product's own enable logic): s1_exec[9], d1_exec[9*noab], d2_exec[9*nvab].
void oracle_task_exec(int noab, int nvab, const int* tiles, const int* spins, int is_restricted,
                      const long* task, int* s1_exec, int* d1_exec, int* d2_exec) {
  SOProblem P = make_problem(noab, nvab, tiles, spins, nullptr, nullptr, nullptr, nullptr, nullptr,
                             nullptr);
  std::vector<int> s1_size(6), d1_size(7 * noab), d2_size(7 * nvab);
  std::fill(s1_exec, s1_exec + 9, -1);
  std::fill(d1_exec, d1_exec + 9 * noab, -1);
  std::fill(d2_exec, d2_exec + 9 * nvab, -1);
  int n1, n2, n3;
  size_t t[6];
  for(int i = 0; i < 6; i++) t[i] = (size_t) task[i];
  data_s1_new(is_restricted != 0, P, t[0], t[1], t[2], t[3], t[4], t[5], 9, 9, s1_size.data(),
              s1_exec, nullptr, nullptr, &n1, false, nullptr);
  data_d1_new(is_restricted != 0, P, t[0], t[1], t[2], t[3], t[4], t[5], 1, 1, 1, nullptr, nullptr,
              nullptr, d1_size.data(), d1_exec, &n2, false, nullptr);
  data_d2_new(is_restricted != 0, P, t[0], t[1], t[2], t[3], t[4], t[5], 1, 1, 1, nullptr, nullptr,
              nullptr, d2_size.data(), d2_exec, &n3, false, nullptr);
}

// Spin-orbital expansion of closed-shell spatial inputs (Appendix of the
// reference's input construction):
//   v2ijab(h1,h2,p1,p2) = (h1 p1|h2 p2) - (h1 p2|h2 p1)       v2tensors.cpp:67-70
//   v2ijka(h1,h2,h3,p1) = (h1 h3|h2 p1) - (h2 h3|h1 p1)       v2tensors.cpp:76-78
//   v2iabc(h1,p1,p2,p3) = (h1 p2|p1 p3) - (h1 p3|p1 p2)       v2tensors.cpp:84-86
// with spin-orbital (pq|rs) = (pq|rs)_spatial [s_p=s_q][s_r=s_s] (Cholesky
// vectors are spin-diagonal).  T1 is spin-diagonal, T2 blocks follow
// cd_ccsd_cs_ann.cpp:19-35 (aaaa = abab - abab(ba)) and 960-976
// (baba, abba, baab), i.e. t2(a,b,i,j) = [s_a=s_i][s_b=s_j] t_ij^ab - [s_a=s_j][s_b=s_i] t_ji^ab.
// Spatial inputs (row-major): t1s[V][O] = t_i^a, t2s[V][V][O][O] = t_ij^ab,
// vvoo[V][V][O][O] = (ai|bj), ovoo[O][V][O][O] = (ia|jk), vvvo[V][V][V][O] = (ab|ci).
// Spin-orbital order is [occ alpha, occ beta] and [virt alpha, virt beta].
void oracle_expand_closed_shell(int O, int V, const double* t1s, const double* t2s,
                                const double* vvoo, const double* ovoo, const double* vvvo,
                                double* t1, double* t2, double* v2ijab, double* v2ijka,
                                double* v2iabc) {
  SpatialSrc S;
  S.O = O;
  S.V = V;
  S.t1s = t1s;
  S.t2s = t2s;
  S.vvoo = vvoo;
  S.ovoo = ovoo;
  S.vvvo = vvvo;
  const long nO = 2L * O, nV = 2L * V;
  for(long a = 0; a < nV; a++)
    for(long i = 0; i < nO; i++) {
      long q[2] = {a, i};
      t1[a * nO + i] = so_elem(S, K_T1, q);
    }
#pragma omp parallel for collapse(2)
  for(long a = 0; a < nV; a++)
    for(long b = 0; b < nV; b++)
      for(long i = 0; i < nO; i++)
        for(long j = 0; j < nO; j++) {
          long q[4] = {a, b, i, j};
          t2[((a * nV + b) * nO + i) * nO + j] = so_elem(S, K_T2, q);
        }
#pragma omp parallel for collapse(2)
  for(long i = 0; i < nO; i++)
    for(long j = 0; j < nO; j++)
      for(long a = 0; a < nV; a++)
        for(long b = 0; b < nV; b++) {
          long q[4] = {i, j, a, b};
          v2ijab[((i * nO + j) * nV + a) * nV + b] = so_elem(S, K_IJAB, q);
        }
#pragma omp parallel for collapse(2)
  for(long i = 0; i < nO; i++)
    for(long j = 0; j < nO; j++)
      for(long k = 0; k < nO; k++)
        for(long a = 0; a < nV; a++) {
          long q[4] = {i, j, k, a};
          v2ijka[((i * nO + j) * nO + k) * nV + a] = so_elem(S, K_IJKA, q);
        }
#pragma omp parallel for collapse(2)
  for(long i = 0; i < nO; i++)
    for(long a = 0; a < nV; a++)
      for(long b = 0; b < nV; b++)
        for(long c = 0; c < nV; c++) {
          long q[4] = {i, a, b, c};
          v2iabc[((i * nV + a) * nV + b) * nV + c] = so_elem(S, K_IABC, q);
        }
}

// Same as oracle_run, with spin-orbital elements formed on the fly from the
// closed-shell spatial inputs (no dense spin-orbital arrays; used where those
// would not fit in memory).  Tiling must mirror alpha/beta with O, V per spin.
int oracle_run_spatial(int noab, int nvab, const int* tiles, const int* spins, const double* evl,
                       int is_restricted, int O, int V, const double* t1s, const double* t2s,
                       const double* vvoo, const double* ovoo, const double* vvvo, long nwhich,
                       const long* which, double* e_out, double* per_task);

// Full fused (T) on the CPU: the reference's task loop over all tasks (or the
// task subset `which`), tasks in driver order.  e_out[2] = energy_l after the
// loop; per_task[2*n] = each task's contribution (factor included).
static int run_impl(SOProblem& P, int is_restricted, long nwhich, const long* which,
                    double* e_out, double* per_task);

int oracle_run(int noab, int nvab, const int* tiles, const int* spins, const double* evl,
               int is_restricted, const double* t1, const double* t2, const double* v2ijab,
               const double* v2ijka, const double* v2iabc, long nwhich, const long* which,
               double* e_out, double* per_task) {
  SOProblem P = make_problem(noab, nvab, tiles, spins, evl, t1, t2, v2ijab, v2ijka, v2iabc);
  return run_impl(P, is_restricted, nwhich, which, e_out, per_task);
}

int oracle_run_spatial(int noab, int nvab, const int* tiles, const int* spins, const double* evl,
                       int is_restricted, int O, int V, const double* t1s, const double* t2s,
                       const double* vvoo, const double* ovoo, const double* vvvo, long nwhich,
                       const long* which, double* e_out, double* per_task) {
  SpatialSrc S;
  S.O = O;
  S.V = V;
  S.t1s = t1s;
  S.t2s = t2s;
  S.vvoo = vvoo;
  S.ovoo = ovoo;
  S.vvvo = vvvo;
  SOProblem P = make_problem(noab, nvab, tiles, spins, evl, nullptr, nullptr, nullptr, nullptr,
                             nullptr);
  if(P.nO != 2 * (size_t) O || P.nV != 2 * (size_t) V) return -2;
  P.src = &S;
  return run_impl(P, is_restricted, nwhich, which, e_out, per_task);
}

// Synthetic inputs of the counter-based generator as arrays (any output may
// be null): the same layouts and values as the product's xt_synth_closed_shell.
void oracle_gen_synth(int O, int V, uint64_t seed, double sigma, double t1_sigma, double* eo,
                      double* ev, double* t1, double* t2, double* vvoo, double* ovoo,
                      double* vvvo) {
  const GenParams G(O, V, seed, sigma, t1_sigma);
  const long      o = O, v = V;
  if(eo) std::copy(G.eo.begin(), G.eo.end(), eo);
  if(ev) std::copy(G.ev.begin(), G.ev.end(), ev);
  if(t1)
    for(long a = 0; a < v; a++)
      for(long i = 0; i < o; i++) t1[a * o + i] = G.t1(a, i);
#pragma omp parallel for
  for(long a = 0; a < v; a++)
    for(long b = 0; b < v; b++)
      for(long i = 0; i < o; i++)
        for(long j = 0; j < o; j++) {
          const size_t id = ((a * v + b) * o + i) * o + j;
          if(vvoo) vvoo[id] = G.vvoo(a, b, i, j);
          if(t2) t2[id] = G.t2(a, b, i, j);
        }
  if(ovoo)
#pragma omp parallel for
    for(long i = 0; i < o; i++)
      for(long a = 0; a < v; a++)
        for(long j = 0; j < o; j++)
          for(long k = 0; k < o; k++) ovoo[((i * v + a) * o + j) * o + k] = G.ovoo(i, a, j, k);
  if(vvvo)
#pragma omp parallel for
    for(long a = 0; a < v; a++)
      for(long b = 0; b < v; b++)
        for(long c = 0; c < v; c++)
          for(long i = 0; i < o; i++) vvvo[((a * v + b) * v + c) * o + i] = G.vvvo(a, b, c, i);
}

// oracle_run_spatial on the generator's inputs, each element formed where the
// gathers read it, so no input array is held (orbital energies in tile order
// |occ_a|occ_b|virt_a|virt_b| from the generator's eo, ev).
int oracle_run_generated(int noab, int nvab, const int* tiles, const int* spins,
                         int is_restricted, int O, int V, uint64_t seed, double sigma,
                         double t1_sigma, long nwhich, const long* which, double* e_out,
                         double* per_task) {
  const GenParams     G(O, V, seed, sigma, t1_sigma);
  std::vector<double> evl;
  for(int s = 0; s < 2; s++) evl.insert(evl.end(), G.eo.begin(), G.eo.end());
  for(int s = 0; s < 2; s++) evl.insert(evl.end(), G.ev.begin(), G.ev.end());
  SpatialSrc S;
  S.O   = O;
  S.V   = V;
  S.gen = &G;
  SOProblem P = make_problem(noab, nvab, tiles, spins, evl.data(), nullptr, nullptr, nullptr,
                             nullptr, nullptr);
  if(P.nO != 2 * (size_t) O || P.nV != 2 * (size_t) V) return -2;
  P.src = &S;
  return run_impl(P, is_restricted, nwhich, which, e_out, per_task);
}

static int run_impl(SOProblem& P, int is_restricted, long nwhich, const long* which,
                    double* e_out, double* per_task) {
  std::vector<size_t> tasks;
  std::vector<double> factors;
  build_tasks(P, is_restricted != 0, tasks, factors);
  Sizes       S = driver_sizes(P);
  TaskBuffers B;
  long double energy_l[2] = {0.0L, 0.0L};  // the reference's energy_l, carried in 80 bits
  long        n = nwhich > 0 ? nwhich : (long) factors.size();
  for(long q = 0; q < n; q++) {
    long k = nwhich > 0 ? which[q] : q;
    if(k < 0 || k >= (long) factors.size()) return -1;
    const size_t* t = &tasks[6 * k];
    gather_task(is_restricted != 0, P, S, t, B);
    double e[2];
    task_kernel_cpu(P, S, t, factors[k], B, e);
    energy_l[0] += e[0];
    energy_l[1] += e[1];
    if(per_task) {
      per_task[2 * q]     = e[0];
      per_task[2 * q + 1] = e[1];
    }
  }
  e_out[0] = (double) energy_l[0];
  e_out[1] = (double) energy_l[1];
  return 0;
}

// Maximum packed-buffer sizes the fused driver allocates (doubles):
// sizes_out = {s1_t1, s1_v2, d1_t2, d1_v2, d2_t2, d2_v2, max_d1_kernels, max_d2_kernels}.
void oracle_buffer_sizes(int noab, int nvab, const int* tiles, const int* spins, long* sizes_out) {
  SOProblem P = make_problem(noab, nvab, tiles, spins, nullptr, nullptr, nullptr, nullptr, nullptr,
                             nullptr);
  Sizes S = driver_sizes(P);
  long v[8] = {(long) S.s1_t1, (long) S.s1_v2, (long) S.d1_t2, (long) S.d1_v2,
               (long) S.d2_t2, (long) S.d2_v2, (long) S.max_d1_kernels_pertask,
               (long) S.max_d2_kernels_pertask};
  for(int i = 0; i < 8; i++) sizes_out[i] = v[i];
}

// The packed buffers and metadata one task hands to the GPU launcher
// (ccsd_t_all_fused.hpp:144-213): buffers sized by oracle_buffer_sizes,
// meta ints: s1_size[6], s1_exec[9], d1_size[7*noab], d1_exec[9*noab],
// d2_size[7*nvab], d2_exec[9*nvab]; counts[3] = enabled s1/d1/d2 buffers.
int oracle_task_packed(int noab, int nvab, const int* tiles, const int* spins, int is_restricted,
                       const double* t1, const double* t2, const double* v2ijab,
                       const double* v2ijka, const double* v2iabc, const long* task,
                       double* s1_t1, double* s1_v2, double* d1_t2, double* d1_v2, double* d2_t2,
                       double* d2_v2, int* meta, int* counts) {
  SOProblem P = make_problem(noab, nvab, tiles, spins, nullptr, t1, t2, v2ijab, v2ijka, v2iabc);
  Sizes       S = driver_sizes(P);
  TaskBuffers B;
  size_t      t[6];
  for(int i = 0; i < 6; i++) t[i] = (size_t) task[i];
  gather_task(is_restricted != 0, P, S, t, B);
  std::copy(B.s1_t1.begin(), B.s1_t1.end(), s1_t1);
  std::copy(B.s1_v2.begin(), B.s1_v2.end(), s1_v2);
  std::copy(B.d1_t2.begin(), B.d1_t2.end(), d1_t2);
  std::copy(B.d1_v2.begin(), B.d1_v2.end(), d1_v2);
  std::copy(B.d2_t2.begin(), B.d2_t2.end(), d2_t2);
  std::copy(B.d2_v2.begin(), B.d2_v2.end(), d2_v2);
  int* m = meta;
  for(int v: B.s1_size) *m++ = v;
  for(int v: B.s1_exec) *m++ = v;
  for(int v: B.d1_size) *m++ = v;
  for(int v: B.d1_exec) *m++ = v;
  for(int v: B.d2_size) *m++ = v;
  for(int v: B.d2_exec) *m++ = v;
  counts[0] = B.n_s1;
  counts[1] = B.n_d1;
  counts[2] = B.n_d2;
  return 0;
}

// The three host gathers of one task from the spatial inputs, written into
// caller-allocated buffers sized by oracle_buffer_sizes without clearing them
// first, as the reference's driver reuses its buffers
// (ccsd_t_all_fused.hpp:149-178).  Returns the doubles written (the enabled
// blocks); meta/counts as oracle_task_packed.  For timing the host gather.
long oracle_gather_spatial(int noab, int nvab, const int* tiles, const int* spins, int is_restricted,
                           int O, int V, const double* t1s, const double* t2s, const double* vvoo,
                           const double* ovoo, const double* vvvo, const long* task, double* s1_t1,
                           double* s1_v2, double* d1_t2, double* d1_v2, double* d2_t2,
                           double* d2_v2, int* meta, int* counts) {
  SpatialSrc S;
  S.O = O;
  S.V = V;
  S.t1s = t1s;
  S.t2s = t2s;
  S.vvoo = vvoo;
  S.ovoo = ovoo;
  S.vvvo = vvvo;
  SOProblem P = make_problem(noab, nvab, tiles, spins, nullptr, nullptr, nullptr, nullptr, nullptr,
                             nullptr);
  if(P.nO != 2 * (size_t) O || P.nV != 2 * (size_t) V) return -2;
  P.src = &S;
  const Sizes Z = driver_sizes(P);
  size_t      t[6];
  for(int i = 0; i < 6; i++) t[i] = (size_t) task[i];
  int* s1_size = meta;
  int* s1_exec = s1_size + 6;
  int* d1_size = s1_exec + 9;
  int* d1_exec = d1_size + 7 * noab;
  int* d2_size = d1_exec + 9 * noab;
  int* d2_exec = d2_size + 7 * nvab;
  std::fill(s1_exec, s1_exec + 9, -1);
  std::fill(d1_exec, d1_exec + 9 * noab, -1);
  std::fill(d2_exec, d2_exec + 9 * nvab, -1);
  size_t elems = 0;
  data_s1_new(is_restricted != 0, P, t[0], t[1], t[2], t[3], t[4], t[5], Z.s1_t1, Z.s1_v2, s1_size,
              s1_exec, s1_t1, s1_v2, &counts[0], true, &elems);
  data_d1_new(is_restricted != 0, P, t[0], t[1], t[2], t[3], t[4], t[5], Z.max_d1_kernels_pertask,
              Z.d1_t2, Z.d1_v2, d1_t2, d1_v2, nullptr, d1_size, d1_exec, &counts[1], true, &elems);
  data_d2_new(is_restricted != 0, P, t[0], t[1], t[2], t[3], t[4], t[5], Z.max_d2_kernels_pertask,
              Z.d2_t2, Z.d2_v2, d2_t2, d2_v2, nullptr, d2_size, d2_exec, &counts[2], true, &elems);
  return (long) elems;
}

int oracle_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// Threads of the calling thread's later parallel regions (OpenMP's per-thread
// setting, so a host thread can time the oracle beside other work).
void oracle_set_num_threads(int n) {
#ifdef _OPENMP
  if(n > 0) omp_set_num_threads(n);
#else
  (void) n;
#endif
}

}  // extern "C"
