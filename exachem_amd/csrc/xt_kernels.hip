// gfx950 kernels of the closed-shell CCSD(T) triples energy.
//
// Work decomposition (shared by the HBM-resident path and the per-task
// packed-buffer drop-in):
//   * a task (h1b,h2b,h3b,p4b,p5b,p6b) is cut into 4x4x4x4x4x4 sub-tiles of
//     the triples tile t3[h1,h2,h3,p4,p5,p6]; one 256-thread workgroup (4
//     wave64) owns one sub-tile at a time and keeps it in LDS (never in HBM);
//   * every d1 and d2 equation of the reference (ccsd_t_all_fused_cpu.hpp:
//     173-270, 331-428) is a GEMM into one of 9 bipartitions of the six
//     t3 indices: M = {single hole hs, particle pair (pA,pB)} and
//     N = {hole pair (hA,hB), single particle ps}.  Each bipartition g is
//     fed by exactly one d1 equation (K = h7) and one d2 equation (K = p7),
//     so the two run back to back into one 64x64 FP64 accumulator held in
//     MFMA registers (v_mfma_f64_16x16x4f64, 2x2 tiles per wave) and the
//     result is added once into the LDS t3 tile;
//   * operands are staged through LDS in K-chunks of 16 (aligned 16-byte
//     loads from zero-padded arrays) with a register prefetch of the next chunk;
//     a workgroup holds 53 KB of LDS, so three share a CU;
//   * s1 (ccsd_t_all_fused_cpu.hpp:472-578) and the denominator-weighted
//     energy (586-619) run on the LDS tile; the 256 partials are reduced
//     with wave64 shuffles and a fixed-order cross-wave sum.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "xt_device.h"
#include "xt_fixed.h"
#include "xt_queue.h"
#include "xt_tasksums.h"

namespace xt {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int NT      = 256;       // threads per workgroup
constexpr int KC      = 16;        // K chunk staged per LDS round
constexpr int LDK     = KC + 2;    // padded row stride (doubles) of staged operands: 144 B, so
                                   // the 16x16 fragment reads (16 rows x 2 k per half-wave) hit
                                   // distinct banks
constexpr int T3N     = 4096;      // 4^6 sub-tile
constexpr int S1_T1   = 9 * 16;    // s1 T1 blocks in LDS
constexpr int STAGE_D = 2 * 64 * LDK;  // staging union: A,B chunks; then the s1 V blocks (9*256)
static_assert(STAGE_D >= 9 * 256, "staging too small for the s1 V blocks");

// LDS of the persistent FP64 kernel, sized for three workgroups per CU.
struct __align__(16) Smem {
  double t3[T3N];
  double stage[STAGE_D];
  double t1s[S1_T1];  // s1 T1 blocks
  double eps[24];
  double red[8];
  int    sub[6], dim[6], off[6];  // current sub-tile: position, extents, offsets
};

__device__ __forceinline__ int cdiv(int a, int b) { return (a + b - 1) / b; }
// parity sign of the single index at position x (0,1,2) within its triple:
// (x, others ascending) is even for x=0,2 and odd for x=1.
__device__ __forceinline__ double sgn_single(int x) { return x == 1 ? -1.0 : 1.0; }

__device__ __forceinline__ void pair_of(int s, int& a, int& b) {
  a = (s == 0) ? 1 : 0;
  b = (s == 2) ? 1 : 2;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for(int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum of a per-thread value over the workgroup in a fixed order; result valid
// in thread 0.
template<class SM>
__device__ __forceinline__ void block_sum2(SM& sm, double& a, double& b) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  a = wave_sum(a);
  b = wave_sum(b);
  if(lane == 0) {
    sm.red[wave]     = a;
    sm.red[4 + wave] = b;
  }
  __syncthreads();
  if(threadIdx.x == 0) {
    a = ((sm.red[0] + sm.red[1]) + sm.red[2]) + sm.red[3];
    b = ((sm.red[4] + sm.red[5]) + sm.red[6]) + sm.red[7];
  }
}

// Index of t3 element (m, n) of bipartition g in the canonical LDS layout
// e = sum_pos idx[pos] << 2*pos (pos 0..2 = h1..h3, 3..5 = p4..p6).
__device__ __forceinline__ int emap(int hs, int hA, int hB, int ps, int pA, int pB, int m, int n) {
  return ((m & 3) << (2 * hs)) | (((m >> 2) & 3) << (2 * (3 + pA))) | ((m >> 4) << (2 * (3 + pB))) |
         ((n & 3) << (2 * hA)) | (((n >> 2) & 3) << (2 * hB)) | ((n >> 4) << (2 * (3 + ps)));
}

// The 64x64xKC MFMA step of one staged chunk: wave w computes rows
// 32*(w>>1).. and columns 32*(w&1).. as 2x2 16x16 tiles.
template<int KCT = KC, int LDKT = LDK>
__device__ __forceinline__ void mfma_chunk(const double* __restrict__ As,
                                           const double* __restrict__ Bs, v4d (&acc)[2][2]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rb = (wave >> 1) * 32, cb = (wave & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
  for(int kk = 0; kk < KCT / 4; kk++) {
    const int k  = 4 * kk + lk;
    double    a0 = As[(rb + lr) * LDKT + k];
    double    a1 = As[(rb + 16 + lr) * LDKT + k];
    double    b0 = Bs[(cb + lr) * LDKT + k];
    double    b1 = Bs[(cb + 16 + lr) * LDKT + k];
    acc[0][0]    = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1]    = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0]    = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1]    = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
  }
}

// Adds the accumulator of bipartition g into the LDS t3 tile.  D layout of
// v_mfma_f64_16x16x4f64: register r of lane l holds row 4*r + (l>>4), column
// l&15 (the four row groups are interleaved, unlike the 16-bit forms).
template<class SM>
__device__ __forceinline__ void acc_to_t3(SM& sm, int g, const v4d (&acc)[2][2]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rb = (wave >> 1) * 32, cb = (wave & 1) * 32;
  const int hs = g / 3, ps = g % 3;
  int       hA, hB, pA, pB;
  pair_of(hs, hA, hB);
  pair_of(ps, pA, pB);
#pragma unroll
  for(int i = 0; i < 2; i++)
#pragma unroll
    for(int j = 0; j < 2; j++)
#pragma unroll
      for(int r = 0; r < 4; r++) {
        const int m = rb + 16 * i + 4 * r + (lane >> 4);
        const int n = cb + 16 * j + (lane & 15);
        sm.t3[emap(hs, hA, hB, ps, pA, pB, m, n)] += acc[i][j][r];
      }
}

// Energy of the finished sub-tile: t3d in LDS, s1 operands staged in LDS,
// denominators from sm.eps.  Returns (E4, E5) partial sums in thread 0.
//   T1s[s][pa][hi] (16 per slot), V1s[s][hk + 4 hj + 16 pc + 64 pb] (256 per slot)
template<class SM>
__device__ __forceinline__ void subtile_energy(SM& sm, const double* T1s, const double* V1s,
                                               const int* sub, const int* dim, unsigned s1mask,
                                               double& e4, double& e5) {
  e4 = 0.0;
  e5 = 0.0;
#pragma unroll 1
  for(int r = 0; r < T3N / NT; r++) {
    const int e = threadIdx.x + NT * r;
    int       idx[6];
    bool      ok = true;
#pragma unroll
    for(int p = 0; p < 6; p++) {
      idx[p] = (e >> (2 * p)) & 3;
      ok     = ok && (4 * sub[p] + idx[p] < dim[p]);
    }
    if(!ok) continue;
    const double d = sm.t3[e];
    double       s = 0.0;
#pragma unroll
    for(int slot = 0; slot < 9; slot++) {
      if(!((s1mask >> slot) & 1)) continue;
      const int pa = slot / 3, hi = slot % 3;
      int       hj, hk, pb, pc;
      pair_of(hi, hj, hk);
      pair_of(pa, pb, pc);
      s += T1s[slot * 16 + idx[3 + pa] * 4 + idx[hi]] *
           V1s[slot * 256 + idx[hk] + 4 * idx[hj] + 16 * idx[3 + pc] + 64 * idx[3 + pb]];
    }
    // denominator in the reference's order (ccsd_t_all_fused_cpu.hpp:612-614)
    const double D = sm.eps[8 + idx[2]] + sm.eps[4 + idx[1]] + sm.eps[idx[0]] -
                     sm.eps[20 + idx[5]] - sm.eps[16 + idx[4]] - sm.eps[12 + idx[3]];
    e4 += d * d / D;
    e5 += d * (d + s) / D;
  }
  block_sum2(sm, e4, e5);
}

// ---------------------------------------------------------------------------
// HBM-resident path: operands are formed on the fly from spatial closed-shell
// tensors (spin-orbital antisymmetrised values of <pq||rs> and T2, exactly the
// numbers the reference's gathers would copy out of TAMM blocks).
// ---------------------------------------------------------------------------
// One staged operand row: pointer to its contraction run and the coefficient
// (+-1 or 0) that folds the equation sign and the spin-orbital selection.
struct RowPtr {
  const double* p;
  double        c;
};

// Four consecutive k of a row (k0 a multiple of 4; rows are padded to a
// multiple of 8, so the run stays inside the row).
__device__ __forceinline__ void fetch4(const RowPtr& rp, int k0, int Kp, double* x) {
  if(rp.c != 0.0 && k0 < Kp) {
    const double2* q  = reinterpret_cast<const double2*>(rp.p + k0);
    double2        v0 = q[0], v1 = q[1];
    x[0] = v0.x; x[1] = v0.y; x[2] = v1.x; x[3] = v1.y;
  }
  else {
#pragma unroll
    for(int j = 0; j < 4; j++) x[j] = 0.0;
  }
}

__device__ void spatial_subtile(Smem& sm, const SpatialDev& S, uint64_t d1code, uint64_t d2code,
                                uint64_t s1code, double& e4, double& e5) {
  const int  tid = threadIdx.x;
  const int  O = S.O, V = S.V;
  const int* sub = sm.sub;
  for(int e = tid; e < T3N; e += NT) sm.t3[e] = 0.0;
  if(tid < 24) {
    const int p = tid >> 2, i = tid & 3, g = 4 * sub[p] + i;
    sm.eps[tid] = g < sm.dim[p] ? (p < 3 ? S.eo : S.ev)[sm.off[p] + g] : 0.0;
  }
  __syncthreads();

  const int row = tid >> 2, kq = tid & 3;
  double*   As  = sm.stage;
  double*   Bs  = sm.stage + 64 * LDK;

#pragma unroll 1
  for(int g = 0; g < 9; g++) {
    const unsigned c1 = (unsigned) (d1code >> (6 * g)) & 63u;
    const unsigned c2 = (unsigned) (d2code >> (6 * g)) & 63u;
    const int      n1 = (c1 & 1) ? cdiv(O, KC) : 0;
    const int      n2 = (c2 & 1) ? cdiv(V, KC) : 0;
    if(n1 + n2 == 0) continue;
    const int hs = g / 3, ps = g % 3;
    int       hA, hB, pA, pB;
    pair_of(hs, hA, hB);
    pair_of(ps, pA, pB);
    // this thread's M row (hs, pA, pB) and N row (hA, hB, ps)
    const int gHS = 4 * sub[hs] + (row & 3), gPA = 4 * sub[3 + pA] + ((row >> 2) & 3),
              gPB = 4 * sub[3 + pB] + (row >> 4);
    const int gHA = 4 * sub[hA] + (row & 3), gHB = 4 * sub[hB] + ((row >> 2) & 3),
              gPS = 4 * sub[3 + ps] + (row >> 4);
    const bool mok = gHS < sm.dim[hs] && gPA < sm.dim[3 + pA] && gPB < sm.dim[3 + pB];
    const bool nok = gHA < sm.dim[hA] && gHB < sm.dim[hB] && gPS < sm.dim[3 + ps];
    const size_t HS = sm.off[hs] + (mok ? gHS : 0), PA = sm.off[3 + pA] + (mok ? gPA : 0),
                 PB = sm.off[3 + pB] + (mok ? gPB : 0);
    const size_t HA = sm.off[hA] + (nok ? gHA : 0), HB = sm.off[hB] + (nok ? gHB : 0),
                 PS = sm.off[3 + ps] + (nok ? gPS : 0);
    // operand rows c1*x - c2*y of the spin-orbital elements: both terms ->
    // the antisymmetrised array; one term -> that array (sign folded into c)
    auto pick = [](bool ok, unsigned code, unsigned bx, unsigned by, double sgn, const double* both,
                   const double* x, const double* y) {
      RowPtr r;
      const bool hx = (code & bx) != 0, hy = (code & by) != 0;
      r.p = hx && hy ? both : (hx ? x : y);
      r.c = !ok || !(hx || hy) ? 0.0 : (hx ? sgn : -sgn);
      return r;
    };
    const size_t Op = S.Op, Vp = S.Vp;
    RowPtr       a1r, b1r, a2r, b2r;
    {
      const double sg = (c1 & 2) ? -1.0 : 1.0;
      const size_t xo = ((PA * V + PB) * O + HS) * Op, yo = ((PB * V + PA) * O + HS) * Op;
      // t(pA,pB,hs,h7) - t(pA,pB,h7,hs) = t2h[pA,pB,hs,:] - t2h[pB,pA,hs,:]
      a1r = pick(mok, c1, 4, 8, sg, S.t2ha + xo, S.t2h + xo, S.t2h + yo);
      // (hA h7|hB ps) - (hB h7|hA ps) = ovk[hB,ps,hA,:] - ovk[hA,ps,hB,:]
      const size_t bxo = ((HB * V + PS) * O + HA) * Op, byo = ((HA * V + PS) * O + HB) * Op;
      b1r = pick(nok, c1, 16, 32, 1.0, S.ovka + bxo, S.ovk + bxo, S.ovk + byo);
    }
    {
      const double sg = (c2 & 2) ? -1.0 : 1.0;
      // (hs pA|p7 pB) - (hs pB|p7 pA) = xv[pB,pA,hs,:] - xv[pA,pB,hs,:]
      const size_t xo = ((PB * V + PA) * O + HS) * Vp, yo = ((PA * V + PB) * O + HS) * Vp;
      a2r = pick(mok, c2, 4, 8, sg, S.xva + xo, S.xv + xo, S.xv + yo);
      // t(ps,p7,hA,hB) - t(p7,ps,hA,hB) = t2p[hA,hB,ps,:] - t2p[hB,hA,ps,:]
      const size_t bxo = ((HA * O + HB) * V + PS) * Vp, byo = ((HB * O + HA) * V + PS) * Vp;
      b2r = pick(nok, c2, 16, 32, 1.0, S.t2pa + bxo, S.t2p + bxo, S.t2p + byo);
    }

    v4d acc[2][2];
#pragma unroll
    for(int i = 0; i < 2; i++)
#pragma unroll
      for(int j = 0; j < 2; j++) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

    double    ax[4], bx[4];
    const int nc    = n1 + n2;
    auto      fetch = [&](int c) {
      if(c < n1) {
        fetch4(a1r, c * KC + 4 * kq, (int) Op, ax);
        fetch4(b1r, c * KC + 4 * kq, (int) Op, bx);
      }
      else {
        fetch4(a2r, (c - n1) * KC + 4 * kq, (int) Vp, ax);
        fetch4(b2r, (c - n1) * KC + 4 * kq, (int) Vp, bx);
      }
    };
    fetch(0);
#pragma unroll 1
    for(int c = 0; c < nc; c++) {
      const double ca = c < n1 ? a1r.c : a2r.c, cb = c < n1 ? b1r.c : b2r.c;
      __syncthreads();
#pragma unroll
      for(int j = 0; j < 4; j++) {
        As[row * LDK + 4 * kq + j] = ca * ax[j];
        Bs[row * LDK + 4 * kq + j] = cb * bx[j];
      }
      __syncthreads();
      if(c + 1 < nc) fetch(c + 1);
      mfma_chunk(As, Bs, acc);
    }
    acc_to_t3(sm, g, acc);
  }

  // ---- s1 operands: T1so(pa,hi) and <hk hj||pc pb> blocks of the sub-tile
  __syncthreads();
  double*  T1s    = sm.t1s;
  double*  V1s    = sm.stage;
  unsigned s1mask = 0;
#pragma unroll
  for(int slot = 0; slot < 9; slot++)
    if((s1code >> (4 * slot)) & 1) s1mask |= 1u << slot;
  for(int x = tid; x < 9 * 16; x += NT) {
    const int      slot = x >> 4, pai = (x >> 2) & 3, hii = x & 3;
    const unsigned c    = (unsigned) (s1code >> (4 * slot)) & 15u;
    const int      pa = slot / 3, hi = slot % 3;
    const int      gp = 4 * sub[3 + pa] + pai, gh = 4 * sub[hi] + hii;
    double         v  = 0.0;
    if((c & 1) && gp < sm.dim[3 + pa] && gh < sm.dim[hi]) {
      v = S.t1[(size_t) (sm.off[3 + pa] + gp) * O + sm.off[hi] + gh];
      if(c & 2) v = -v;
    }
    T1s[x] = v;
  }
  for(int x = tid; x < 9 * 256; x += NT) {
    const int      slot = x >> 8, w = x & 255;
    const unsigned c    = (unsigned) (s1code >> (4 * slot)) & 15u;
    const int      pa = slot / 3, hi = slot % 3;
    int            hj, hk, pb, pc;
    pair_of(hi, hj, hk);
    pair_of(pa, pb, pc);
    const int gk = 4 * sub[hk] + (w & 3), gj = 4 * sub[hj] + ((w >> 2) & 3),
              gc = 4 * sub[3 + pc] + ((w >> 4) & 3), gb = 4 * sub[3 + pb] + (w >> 6);
    double v = 0.0;
    if((c & 1) && gk < sm.dim[hk] && gj < sm.dim[hj] && gc < sm.dim[3 + pc] && gb < sm.dim[3 + pb]) {
      const size_t K = sm.off[hk] + gk, J = sm.off[hj] + gj, PC = sm.off[3 + pc] + gc,
                   PBx = sm.off[3 + pb] + gb;
      const double x1 = (c & 4) ? S.vvoo[((PC * V + PBx) * O + K) * O + J] : 0.0;  // (hk pc|hj pb)
      const double x2 = (c & 8) ? S.vvoo[((PBx * V + PC) * O + K) * O + J] : 0.0;  // (hk pb|hj pc)
      v               = ((c & 4) && (c & 8)) ? x1 - x2 : ((c & 4) ? x1 : -x2);
    }
    V1s[x] = v;
  }
  __syncthreads();
  subtile_energy(sm, T1s, V1s, sm.sub, sm.dim, s1mask, e4, e5);
}

// Persistent kernel: workgroup b serves queue b % kQueues first (dispatch is
// round-robin over the 8 XCDs, so that is its own XCD's queue), then steals
// from the others; one ticket = one sub-tile.  Thread 0 claims and decodes each
// ticket (its entry searched forward from its previous one) and keeps the
// running exact sums of the task its sub-tiles belong to (xt_tasksums.h).
__global__ void __launch_bounds__(NT, 3) k_spatial_fp64(SpatialDev S, RunDev R) {
  __shared__ Smem               sm;
  __shared__ int                s_task;
  __shared__ int                s_entry;   // thread 0: entry of its last ticket on the current queue
  __shared__ int                s_cur;     // thread 0: task of the running sums
  __shared__ unsigned long long s_sum[4];  // thread 0: E4 lo, hi, E5 lo, hi of s_cur not yet added
  const int                     tid  = threadIdx.x;
  const int                     home = blockIdx.x % kQueues;
  if(tid == 0) {
    s_cur    = -1;
    s_sum[0] = s_sum[1] = s_sum[2] = s_sum[3] = 0ull;
  }
#pragma unroll 1
  for(int qi = 0; qi < kQueues; qi++) {
    const int q = (home + qi) % kQueues;
    if(tid == 0) s_entry = R.qbeg[q];
#pragma unroll 1
    for(;;) {
      __syncthreads();  // previous sub-tile finished with LDS and s_task
      if(tid == 0) {
        const unsigned long long t = atomicAdd(&R.qctr[q], 1ull);
        // progress mirror in host memory (tickets taken so far, approximate)
        if(R.prog) __hip_atomic_store(&R.prog[q], t + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        int ti = -1;  // -1: queue exhausted
        if((long long) t < R.qtotal[q]) {
          const int e = queue_entry_from(R.qbeg, R.qtick, q, (long long) t, s_entry);
          s_entry     = e;
          ti          = R.qtask[e];
          const TaskDev& T = R.tasks[ti];
          // ticket -> sub-tile: the low bits select the box inside its group
          int sub[6];
          if(ticket_box((long long) t - R.qtick[e], T.nsub, R.glog, sub)) {
#pragma unroll
            for(int p = 0; p < 6; p++) {
              sm.sub[p] = sub[p];
              sm.dim[p] = T.dim[p];
              sm.off[p] = T.off[p];
            }
          }
          else ti = -2;  // a ticket past the task's sub-tile grid (its last group overhangs)
        }
        s_task = ti;
      }
      __syncthreads();
      const int ti = s_task;
      if(ti == -1) break;
      if(ti == -2) continue;
      const TaskDev&           T  = R.tasks[ti];
      const unsigned long long t0 = wall_clock64();
      double                   e4, e5;
      spatial_subtile(sm, S, T.d1code, T.d2code, T.s1code, e4, e5);
      if(tid == 0) {
        if(ti != s_cur) {
          sums_flush(R, s_cur, s_sum);
          s_cur = ti;
        }
        sums_add(R, s_sum, e4, e5);
        if(R.cost) atomicAdd(&R.cost[ti], wall_clock64() - t0);
      }
    }
  }
  if(tid == 0) sums_flush(R, s_cur, s_sum);
}

// Per-task energy from the fixed-point accumulators, scaled by the task factor.
__global__ void k_reduce_tasks(const TaskDev* tasks, int n_tasks, const unsigned long long* acc,
                               double* task_e) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if(t >= n_tasks) return;
  const double f             = tasks[t].factor;
  task_e[2 * (size_t) t]     = f * fx_value(&acc[4 * (size_t) t]);
  task_e[2 * (size_t) t + 1] = f * fx_value(&acc[4 * (size_t) t + 2]);
}

__global__ void k_zero_u64(unsigned long long* p, long n) {
  const long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) p[i] = 0ull;
}

// Fixed-shape sum over tasks in task order (one workgroup of NT threads):
// thread i sums tasks [i*per, (i+1)*per) in order, per = ceil(n/NT), then a
// halving tree over the threads.
static_assert(NT == kReduceThreads, "the host repeats this reduction with kReduceThreads");
__global__ void __launch_bounds__(NT) k_reduce_final(const double* task_e, int n, double* e45) {
  __shared__ double ra[NT], rb[NT];
  const int         per = (n + NT - 1) / NT;
  const int         lo = threadIdx.x * per, hi = min(n, lo + per);
  double            a = 0.0, b = 0.0;
  for(int t = lo; t < hi; t++) {
    a += task_e[2 * (size_t) t];
    b += task_e[2 * (size_t) t + 1];
  }
  ra[threadIdx.x] = a;
  rb[threadIdx.x] = b;
  __syncthreads();
  for(int s = NT / 2; s >= 1; s >>= 1) {
    if((int) threadIdx.x < s) {
      ra[threadIdx.x] += ra[threadIdx.x + s];
      rb[threadIdx.x] += rb[threadIdx.x + s];
    }
    __syncthreads();
  }
  if(threadIdx.x == 0) {
    e45[0] = ra[0];
    e45[1] = rb[0];
  }
}

// out[b*ldo + a] = in[a*lda + b] for a < rows, b < cols, batched over
// blockIdx.z with strides (in_bs, out_bs).  32x32 LDS tiles.
__global__ void k_transpose(const double* __restrict__ in, double* __restrict__ out, long rows,
                            long cols, long lda, long ldo, long in_bs, long out_bs) {
  __shared__ double tile[32][33];
  const long z  = blockIdx.z;
  const long a0 = (long) blockIdx.y * 32, b0 = (long) blockIdx.x * 32;
  const int  tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 32 x 8
  for(int r = ty; r < 32; r += 8) {
    const long a = a0 + r, b = b0 + tx;
    tile[r][tx]  = (a < rows && b < cols) ? in[z * in_bs + a * lda + b] : 0.0;
  }
  __syncthreads();
  for(int r = ty; r < 32; r += 8) {
    const long b = b0 + r, a = a0 + tx;
    if(a < rows && b < cols) out[z * out_bs + b * ldo + a] = tile[tx][r];
  }
}

static int transpose(const double* in, double* out, long rows, long cols, long lda, long ldo,
                     long batch, long in_bs, long out_bs, hipStream_t st) {
  // grid.y and grid.z are limited to 65535
  const long max_yz = 65535;
  for(long z0 = 0; z0 < batch; z0 += max_yz) {
    const long nz = std::min(max_yz, batch - z0);
    for(long r0 = 0; r0 < rows; r0 += 32 * max_yz) {
      const long nr = std::min(32 * max_yz, rows - r0);
      dim3       grid((unsigned) ((cols + 31) / 32), (unsigned) ((nr + 31) / 32), (unsigned) nz);
      hipLaunchKernelGGL(k_transpose, grid, dim3(256), 0, st, in + z0 * in_bs + r0 * lda,
                         out + z0 * out_bs + r0, nr, cols, lda, ldo, in_bs, out_bs);
      if(hipGetLastError() != hipSuccess) return (int) hipErrorLaunchFailure;
    }
  }
  return 0;
}

// out[r][k] = in[r][k] for k < K, 0 for K <= k < Kp
__global__ void k_pad_rows(double* __restrict__ out, const double* __restrict__ in, long rows,
                           long K, long Kp) {
  const long n = rows * Kp;
  for(long i = (long) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long) gridDim.x * blockDim.x) {
    const long r = i / Kp, k = i % Kp;
    out[i]       = k < K ? in[r * K + k] : 0.0;
  }
}

// out[i0][i1][i2][k] = in[i0][i1][i2][k] - in[j0][j1][j2][k] with (j0,j1,j2) =
// (i1,i0,i2) (swap02 = 0) or (i2,i1,i0) (swap02 = 1); padding stays zero.
__global__ void k_antisym(double* __restrict__ out, const double* __restrict__ in, long n0,
                          long n1, long n2, long Kp, int swap02) {
  const long n = n0 * n1 * n2 * Kp;
  for(long i = (long) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long) gridDim.x * blockDim.x) {
    long       r  = i / Kp;
    const long k  = i % Kp;
    const long i2 = r % n2;
    r /= n2;
    const long i1 = r % n1, i0 = r / n1;
    const long j  = swap02 ? ((i2 * n1 + i1) * n2 + i0) * Kp + k : ((i1 * n1 + i0) * n2 + i2) * Kp + k;
    out[i]        = in[i] - in[j];
  }
}

// Blocks of 256 for n elements of a grid-stride kernel.  A dispatch's grid
// size is a 32-bit count of work-items, so the grid is capped (2^20 blocks
// keep every CU busy) and the kernels loop over the rest.
static unsigned nblk(long n) { return (unsigned) std::min<long>((n + 255) / 256, 1l << 20); }

int build_operands(int O, int V, double* t2, double* ovoo, double*& vvvo, SpatialBufs& B,
                   void* stream) {
  hipStream_t st = (hipStream_t) stream;
  const long  o = O, v = V, op = (O + 7) / 8 * 8, vp = (V + 7) / 8 * 8;
  auto        alloc = [](double** p, long n) { return (hipError_t) track_malloc((void**) p, sizeof(double) * n); };
  hipError_t  e;
  int         rc;
  // d1 operands (contraction over h7, padded to op)
  if((e = alloc(&B.t2h, v * v * o * op))) return (int) e;
  hipLaunchKernelGGL(k_pad_rows, dim3(nblk(v * v * o * op)), dim3(256), 0, st, B.t2h, t2, v * v * o,
                     o, op);
  if((e = alloc(&B.t2ha, v * v * o * op))) return (int) e;
  hipLaunchKernelGGL(k_antisym, dim3(nblk(v * v * o * op)), dim3(256), 0, st, B.t2ha, B.t2h, v, v,
                     o, op, 0);
  if((e = alloc(&B.ovk, o * v * o * op))) return (int) e;
  hipLaunchKernelGGL(k_pad_rows, dim3(nblk(o * v * o * op)), dim3(256), 0, st, B.ovk, ovoo,
                     o * v * o, o, op);
  if((e = alloc(&B.ovka, o * v * o * op))) return (int) e;
  hipLaunchKernelGGL(k_antisym, dim3(nblk(o * v * o * op)), dim3(256), 0, st, B.ovka, B.ovk, o, v,
                     o, op, 1);
  // d2 N side: t2p[i][j][a][b] = t2[a][b][i][j], padded in b
  double* tmp = nullptr;
  if((e = alloc(&tmp, o * o * v * v))) return (int) e;
  if((rc = transpose(t2, tmp, v * v, o * o, o * o, v * v, 1, 0, 0, st))) return rc;
  if((e = alloc(&B.t2p, o * o * v * vp))) return (int) e;
  hipLaunchKernelGGL(k_pad_rows, dim3(nblk(o * o * v * vp)), dim3(256), 0, st, B.t2p, tmp, o * o * v,
                     v, vp);
  if((e = alloc(&B.t2pa, o * o * v * vp))) return (int) e;
  hipLaunchKernelGGL(k_antisym, dim3(nblk(o * o * v * vp)), dim3(256), 0, st, B.t2pa, B.t2p, o, o, v,
                     vp, 0);
  if((e = hipStreamSynchronize(st))) return (int) e;
  track_free(tmp);
  if(!vvvo) return 0;  // xv comes from elsewhere (build_xv_cholesky), then build_xva
  // d2 M side: xv[c][b][i][a] = vvvo[a][c][b][i], padded in a
  if((e = alloc(&B.xv, v * v * o * vp))) return (int) e;
  if((e = hipMemsetAsync(B.xv, 0, sizeof(double) * v * v * o * vp, st))) return (int) e;
  if((rc = transpose(vvvo, B.xv, v, v * o, v * v * o, vp, v, v * o, v * o * vp, st))) return rc;
  if((e = hipStreamSynchronize(st))) return (int) e;
  track_free(vvvo);
  vvvo = nullptr;
  return build_xva(O, V, B, stream);
}

int build_xva(int O, int V, SpatialBufs& B, void* stream) {
  hipStream_t st = (hipStream_t) stream;
  const long  o = O, v = V, vp = (V + 7) / 8 * 8;
  hipError_t  e;
  if((e = (hipError_t) track_malloc((void**) &B.xva, sizeof(double) * v * v * o * vp))) return (int) e;
  hipLaunchKernelGGL(k_antisym, dim3(nblk(v * v * o * vp)), dim3(256), 0, st, B.xva, B.xv, v, v, o,
                     vp, 0);
  if((e = hipGetLastError())) return (int) e;
  return (int) hipStreamSynchronize(st);
}

// out[a][b][i][j] = in[((a * O + i) * V + b) * O + j]: the (ai|bj) GEMM result,
// rows (a,i) and columns (b,j), in the vvoo order.
__global__ void k_vvoo_order(const double* __restrict__ in, double* __restrict__ out, long O, long V) {
  const long n = V * V * O * O;
  for(long x = (long) blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (long) gridDim.x * blockDim.x) {
    const long j = x % O, i = (x / O) % O, b = (x / (O * O)) % V, a = x / (O * O * V);
    out[x]       = in[((a * O + i) * V + b) * O + j];
  }
}

int launch_vvoo_order(const double* in, double* out, int O, int V, void* stream) {
  const long n = (long) V * V * O * O;
  hipLaunchKernelGGL(k_vvoo_order, dim3(nblk(n)), dim3(256), 0, (hipStream_t) stream, in, out,
                     (long) O, (long) V);
  return (int) hipGetLastError();
}

int launch_zero_u64(unsigned long long* p, long n, void* stream) {
  if(n <= 0) return 0;
  hipLaunchKernelGGL(k_zero_u64, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t) stream, p, n);
  return (int) hipGetLastError();
}

int launch_spatial_fp64(const SpatialDev& S, const RunDev& R, int grid, void* stream) {
  hipLaunchKernelGGL(k_spatial_fp64, dim3(grid), dim3(NT), 0, (hipStream_t) stream, S, R);
  return (int) hipGetLastError();
}

int launch_reduce(const TaskDev* tasks, int n_tasks, const unsigned long long* acc,
                  double* task_e, double* e45, void* stream) {
  hipStream_t st = (hipStream_t) stream;
  hipLaunchKernelGGL(k_reduce_tasks, dim3(n_tasks > 0 ? (n_tasks + 255) / 256 : 1), dim3(256), 0, st, tasks,
                     n_tasks, acc, task_e);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(NT), 0, st, task_e, n_tasks, e45);
  return (int) hipGetLastError();
}

// ---------------------------------------------------------------------------
// Per-task drop-in: operands come from the reference's packed buffers
// (ccsd_t_data_{s1,d1,d2}_new layouts, ccsd_t_all_fused_cpu.hpp index
// arithmetic), one workgroup per 4^6 block, partials per block.
// ---------------------------------------------------------------------------
// LDS of the per-task kernel, sized for three workgroups per CU (12 waves):
// the 4^6 tile, one staged K chunk of 16 (rows padded to 18 doubles, so the
// 16x16x4 fragment reads of a half-wave hit distinct banks), which later holds
// the s1 V blocks, and the K tile list of the current bipartition, which
// later holds the s1 T1 blocks.
constexpr int PKC  = 16;
constexpr int PLDK = PKC + 2;
static_assert(PKC == kPackedKC, "the host aligns the per-task d2 tiles to the K chunk");
static_assert(2 * 64 * PLDK >= 9 * 256, "staging too small for the s1 V blocks");
struct __align__(16) PSmem {
  double t3[T3N];
  double stage[2 * 64 * PLDK];
  union {
    int2   klist[kMaxNoab + kMaxNvab + 1];  // (family | exec slot << 1, K - 1 | first k << 12)
    double t1s[S1_T1];
  } u;
  double eps[24];
  double red[8];
};
// a K tile list entry fits two ints: K <= 4096, total K < 2^20, exec slot < 2^30
static_assert((kMaxNoab + kMaxNvab) * 4096 + kPackedKC < (1 << 20), "K offsets must fit 20 bits");

__global__ void __launch_bounds__(NT, 3) k_packed_fp64(PackedArgs A) {
  __shared__ PSmem sm;
  const int        tid = threadIdx.x;
  const int*       dim = A.dim;
  int              nsub[6], sub[6];
#pragma unroll
  for(int p = 0; p < 6; p++) nsub[p] = cdiv(dim[p], 4);
  {
    long q = blockIdx.x;
#pragma unroll
    for(int p = 0; p < 6; p++) {
      sub[p] = (int) (q % nsub[p]);
      q /= nsub[p];
    }
  }
  const int* d1_size = A.meta;
  const int* d1_exec = d1_size + 7 * A.noab;
  const int* d2_size = d1_exec + 9 * A.noab;
  const int* d2_exec = d2_size + 7 * A.nvab;
  const int* s1_exec = d2_exec + 9 * A.nvab + 6;
  // the host's K tile lists of the nine bipartitions (packed_klists)
  const int*  k_n    = s1_exec + 9;
  const int4* k_list = reinterpret_cast<const int4*>(k_n + kKlistCountInts);

  for(int e = tid; e < T3N; e += NT) sm.t3[e] = 0.0;
  if(tid < 24) {
    const int p = tid >> 2, i = tid & 3, g = 4 * sub[p] + i;
    sm.eps[tid] = g < dim[p] ? A.evl[p][g] : 0.0;
  }
  __syncthreads();
  double* As = sm.stage;
  double* Bs = sm.stage + 64 * PLDK;
  // decoded K tile list entry: (family, exec slot, K, first k); first k alone
  auto kent = [&](int i) {
    const int2 v = sm.u.klist[i];
    return make_int4(v.x & 1, v.x >> 1, (v.y & 0xfff) + 1, (int) ((unsigned) v.y >> 12));
  };
  auto kbeg = [&](int i) { return (int) ((unsigned) sm.u.klist[i].y >> 12); };

#pragma unroll 1
  for(int g = 0; g < 9; g++) {
    const int hs = g / 3, ps = g % 3;
    int       hA, hB, pA, pB;
    pair_of(hs, hA, hB);
    pair_of(ps, pA, pB);
    const double sg = -sgn_single(hs) * sgn_single(ps);
    const long   DHS = dim[hs], DPA = dim[3 + pA], DPB = dim[3 + pB];
    const long   DHA = dim[hA], DHB = dim[hB], DPS = dim[3 + ps];
    v4d          acc[2][2];
#pragma unroll
    for(int i = 0; i < 2; i++)
#pragma unroll
      for(int j = 0; j < 2; j++) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
    // K of this bipartition: the enabled d1 tiles (h7b), then the enabled d2
    // tiles (p7b) from the next multiple of PKC, packed end to end and cut
    // into chunks of PKC; each chunk reads one family.  The host lists the
    // tiles (family, exec slot, K, first packed k) with a sentinel holding the
    // total.  Every thread has read the previous bipartition's list before it
    // is rewritten.
    int koff = 0;
    for(int gg = 0; gg < g; gg++) koff += (k_n[gg] & 0xffff) + 1;
    const int nk = k_n[g] & 0xffff, nd1 = k_n[g] >> 16;
    __syncthreads();
    for(int x = tid; x <= nk; x += NT) {
      const int4 e = k_list[koff + x];
      sm.u.klist[x] = make_int2(e.x | (e.y << 1),
                                (int) ((unsigned) (e.z > 0 ? e.z - 1 : 0) | ((unsigned) e.w << 12)));
    }
    __syncthreads();
    const int kall = kbeg(nk), kd2 = kbeg(nd1);  // total K, first d2 k
    // Loads run along k where the packed block keeps k contiguous (d1_t2,
    // d2_v2, d2_t2: row * K + k): thread (kk, rs) = (tid % 16, tid / 16)
    // reads k = kk of rows rs + 16 j, so a load instruction reads four runs of
    // 16 doubles.  The d1 <ij||ka> block keeps k slowest (row + k * bstr1):
    // there thread (n1, kq) = (tid % 64, tid / 64) reads row n1 at k = 4 kq + j,
    // and an instruction reads 64 neighbouring rows.
    // Row r of either operand: r & 3, (r >> 2) & 3, r >> 4 index the first,
    // second and third t3 position of its side (hs, pA, pB or hA, hB, ps), so
    // row rs + 16 j moves the third position by j.
    const int kk = tid & 15, rs = tid >> 4;
    const int aHS = 4 * sub[hs] + (rs & 3), aPA = 4 * sub[3 + pA] + ((rs >> 2) & 3),
              aPB = 4 * sub[3 + pB];
    const int bHA = 4 * sub[hA] + (rs & 3), bHB = 4 * sub[hB] + ((rs >> 2) & 3),
              bPS = 4 * sub[3 + ps];
    unsigned  amask = 0, bmask = 0;  // rows inside the block, per j
#pragma unroll
    for(int j = 0; j < 4; j++) {
      if(aHS < DHS && aPA < DPA && aPB + j < DPB) amask |= 1u << j;
      if(bHA < DHA && bHB < DHB && bPS + j < DPS) bmask |= 1u << j;
    }
    const long a1 = aPA + (aPB + aHS * DPB) * DPA;  // d1_t2 row of j = 0: [hs][pB][pA]
    const long a2 = aHS + (aPB + aPA * DPB) * DHS;  // d2_v2 row: [pA][pB][hs]
    const long b2 = bPS + (bHA + bHB * DHA) * DPS;  // d2_t2 row: [hB][hA][ps]
    const int  n1 = tid & 63, kq = tid >> 6;
    const int  gHA = 4 * sub[hA] + (n1 & 3), gHB = 4 * sub[hB] + ((n1 >> 2) & 3),
              gPS = 4 * sub[3 + ps] + (n1 >> 4);
    const bool nok1  = gHA < DHA && gHB < DHB && gPS < DPS;
    const long rB1   = gHB + (gHA + gPS * DHA) * DHB;  // d1_v2 element of k = rB1 + k * bstr1
    const long bstr1 = DPS * DHA * DHB;
    // tile walkers (k only grows): wa for the thread's k-contiguous k, wb for
    // its d1 <ij||ka> k; te = tile of k, nb = first k of the next tile
    int  tia = 0, tib = 0;
    int4 tea = kent(0), teb = tea;
    int  nba = kbeg(1), nbb = nba;
    double   av[4], bv[4];
    unsigned oka = 0, okb = 0;
    // Issues every load of chunk c (rows outside the block and k outside a
    // tile read the buffer's first element; oka / okb mark the valid ones), so
    // the loads of a chunk are in flight together while the previous chunk is
    // multiplied.
    auto load = [&](int c) {
      const bool d2 = c * PKC >= kd2;
      const int  k  = c * PKC + kk;
      bool       kok = k < kall;
      if(kok)
        while(k >= nba) {
          tea = kent(++tia);
          nba = kbeg(tia + 1);
        }
      const long kl = k - tea.w, K = tea.z, ex = tea.y;
      kok           = kok && kl < K;
      const double* pa[4];
      const double* pb[4];
      {
        const double* p0 = d2 ? A.d2_v2 + A.max_d2_v2 * ex + a2 * K + kl
                              : A.d1_t2 + A.max_d1_t2 * ex + a1 * K + kl;
        const long    s  = d2 ? DHS * K : DPA * K;
        oka              = kok ? amask : 0u;
#pragma unroll
        for(int j = 0; j < 4; j++) pa[j] = ((oka >> j) & 1) ? p0 + j * s : A.d1_t2;
      }
      if(d2) {
        const double* p0 = A.d2_t2 + A.max_d2_t2 * ex + b2 * K + kl;
        okb              = kok ? bmask : 0u;
#pragma unroll
        for(int j = 0; j < 4; j++) pb[j] = ((okb >> j) & 1) ? p0 + j * K : A.d1_v2;
      }
      else {
        okb = 0;
#pragma unroll
        for(int j = 0; j < 4; j++) {
          const int k1 = c * PKC + 4 * kq + j;
          bool      ok = nok1 && k1 < kall;
          if(ok)
            while(k1 >= nbb) {
              teb = kent(++tib);
              nbb = kbeg(tib + 1);
            }
          const long kl1 = k1 - teb.w;
          ok             = ok && kl1 < teb.z;
          pb[j]          = ok ? A.d1_v2 + A.max_d1_v2 * (long) teb.y + rB1 + kl1 * bstr1 : A.d1_v2;
          okb |= (ok ? 1u : 0u) << j;
        }
      }
#pragma unroll
      for(int j = 0; j < 4; j++) {
        av[j] = *pa[j];
        bv[j] = *pb[j];
      }
    };
    if(kall > 0) {
      const int nc = cdiv(kall, PKC);
      load(0);
#pragma unroll 1
      for(int c = 0; c < nc; c++) {
        __syncthreads();
        // the masks are applied here, so that a chunk's loads are waited for
        // only after the previous chunk's MFMAs were issued
#pragma unroll
        for(int j = 0; j < 4; j++) As[(rs + 16 * j) * PLDK + kk] = ((oka >> j) & 1) ? sg * av[j] : 0.0;
        if(c * PKC >= kd2) {
#pragma unroll
          for(int j = 0; j < 4; j++) Bs[(rs + 16 * j) * PLDK + kk] = ((okb >> j) & 1) ? bv[j] : 0.0;
        }
        else {
#pragma unroll
          for(int j = 0; j < 4; j++) Bs[n1 * PLDK + 4 * kq + j] = ((okb >> j) & 1) ? bv[j] : 0.0;
        }
        __syncthreads();
        if(c + 1 < nc) load(c + 1);
        mfma_chunk<PKC, PLDK>(As, Bs, acc);
      }
      acc_to_t3(sm, g, acc);
    }
  }

  // s1 operands from the packed buffers: t1 [hi][pa], v2 [pb][pc][hj][hk]
  __syncthreads();  // every thread is done with the tile list and the staged chunk
  double*  T1s    = sm.u.t1s;
  double*  V1s    = sm.stage;
  unsigned s1mask = 0;
  for(int slot = 0; slot < 9; slot++)
    if(s1_exec[slot] >= 0) s1mask |= 1u << slot;
  for(int x = tid; x < 9 * 16; x += NT) {
    const int slot = x >> 4, pai = (x >> 2) & 3, hii = x & 3;
    const int pa = slot / 3, hi = slot % 3;
    const int gp = 4 * sub[3 + pa] + pai, gh = 4 * sub[hi] + hii;
    double    v  = 0.0;
    if(s1_exec[slot] >= 0 && gp < dim[3 + pa] && gh < dim[hi])
      v = sgn_single(pa) * sgn_single(hi) *
          A.s1_t1[A.max_s1_t1 * (long) s1_exec[slot] + gp + (long) gh * dim[3 + pa]];
    T1s[x] = v;
  }
  for(int x = tid; x < 9 * 256; x += NT) {
    const int slot = x >> 8, w = x & 255;
    const int pa = slot / 3, hi = slot % 3;
    int       hj, hk, pb, pc;
    pair_of(hi, hj, hk);
    pair_of(pa, pb, pc);
    const int gk = 4 * sub[hk] + (w & 3), gj = 4 * sub[hj] + ((w >> 2) & 3),
              gc = 4 * sub[3 + pc] + ((w >> 4) & 3), gb = 4 * sub[3 + pb] + (w >> 6);
    double v = 0.0;
    if(s1_exec[slot] >= 0 && gk < dim[hk] && gj < dim[hj] && gc < dim[3 + pc] && gb < dim[3 + pb])
      v = A.s1_v2[A.max_s1_v2 * (long) s1_exec[slot] + gk +
                  (gj + (gc + (long) gb * dim[3 + pc]) * dim[hj]) * dim[hk]];
    V1s[x] = v;
  }
  __syncthreads();
  double e4, e5;
  subtile_energy(sm, T1s, V1s, sub, dim, s1mask, e4, e5);
  if(tid == 0) {
    A.partial[blockIdx.x]                = e4;
    A.partial[blockIdx.x + A.num_blocks] = e5;
  }
}

int launch_packed_fp64(const PackedArgs& A, void* stream) {
  hipLaunchKernelGGL(k_packed_fp64, dim3((unsigned) A.num_blocks), dim3(NT), 0,
                     (hipStream_t) stream, A);
  return (int) hipGetLastError();
}

// ---------------------------------------------------------------------------
// Allocation-free form of the per-task launch.  The launcher metadata lives in
// `scratch`, one fixed array per GPU for the life of the process.  In host
// code the name `scratch` is only a stand-in for the symbol, so every host use
// goes through the device address hipGetSymbolAddress returns, looked up once
// per device (with that device current) and kept in scratch_ptr.  The host
// arrays reach it through the arguments of k_store_meta launches on the
// caller's stream: ordered with the caller's work, captured by value at
// launch (the caller may reuse its arrays at once), and no allocation, copy
// engine or memset node involved.  Calls that overlap on one GPU race on it.
// ---------------------------------------------------------------------------
static __device__ __align__(256) int scratch[kPackedMetaInts];

namespace {
constexpr int kMetaChunk = 960;  // ints per copy launch (3.8 KB of kernel arguments)
struct MetaChunk {
  int off, n;
  int v[kMetaChunk];
};

__global__ void __launch_bounds__(256) k_store_meta(int* dst, MetaChunk c) {
  for(int i = threadIdx.x; i < c.n; i += blockDim.x) dst[c.off + i] = c.v[i];
}

constexpr int kMaxDevices = 64;
std::mutex    g_scratch_mu;
int*          scratch_ptr[kMaxDevices] = {};

// Device address of this GPU's scratch; nullptr on failure (*err set).
int* scratch_on_current_device(int* err) {
  int dev = 0;
  if((*err = (int) hipGetDevice(&dev))) return nullptr;
  if(dev < 0 || dev >= kMaxDevices) {
    *err = (int) hipErrorInvalidDevice;
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if(!scratch_ptr[dev]) {
    void* p = nullptr;
    if((*err = (int) hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)))) return nullptr;
    scratch_ptr[dev] = static_cast<int*>(p);
  }
  return scratch_ptr[dev];
}
}  // namespace

int launch_packed_fp64_noalloc(PackedArgs& A, const int* host_meta, int n_meta, void* stream,
                               void* after_copy) {
  if(n_meta < 0 || n_meta > kPackedMetaInts) return (int) hipErrorInvalidValue;
  int  err = 0;
  int* dst = scratch_on_current_device(&err);
  if(!dst) return err ? err : (int) hipErrorInvalidSymbol;
  hipStream_t st = (hipStream_t) stream;
  MetaChunk   c  = {};
  for(int off = 0; off < n_meta; off += kMetaChunk) {
    c.off = off;
    c.n   = std::min(kMetaChunk, n_meta - off);
    std::memcpy(c.v, host_meta + off, sizeof(int) * (size_t) c.n);
    hipLaunchKernelGGL(k_store_meta, dim3(1), dim3(256), 0, st, dst, c);
    if((err = (int) hipGetLastError())) return err;
  }
  if(after_copy && (err = (int) hipEventRecord((hipEvent_t) after_copy, st))) return err;
  A.meta = dst;
  return launch_packed_fp64(A, stream);
}

}  // namespace xt
