// Mixed-precision (T) kernel for gfx950.
//
// Same task / bipartition decomposition as the FP64 kernel, with larger
// workgroup boxes and 16-bit matrix cores:
//   * a workgroup owns a (4,4,4 | 8,8,8) box of t3[h1,h2,h3,p4,p5,p6]
//     (32768 elements, FP32 in LDS, never in HBM);
//   * each bipartition g (single hole hs + particle pair | hole pair +
//     single particle ps) is a 256 x 128 GEMM: M = (hs,pA,pB) 4*8*8,
//     N = (hA,hB,ps) 4*4*8, K = h7 (d1) then p7 (d2);
//   * operands are pre-split FP16 pairs y = hi + lo of x * 2^e (per-row
//     power-of-two amplitude scale, max |y| in [2^13, 2^14)); every product is
//     hi*hi + hi*lo + lo*hi on v_mfma_f32_32x32x16_f16 with FP32 accumulation
//     (~22-bit products); the row scales and the spin-orbital sign are undone
//     when a K part ends and the result is added into the FP32 box;
//   * K chunks of 16 stream through LDS with a register prefetch that runs
//     across part and bipartition boundaries; MFMA fragments are double
//     buffered in registers so each chunk's MFMAs overlap the next chunk's
//     LDS reads;
//   * s1, the denominators and the energy partials are FP64, reduced and
//     accumulated exactly like the FP64 kernel.
#include <hip/hip_runtime.h>

#include "xt_device.h"

namespace xt {
namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float    f16v __attribute__((ext_vector_type(16)));

constexpr int NT   = 256;
constexpr int MM   = 256;  // bipartition rows   (hs, pA, pB)
constexpr int NN   = 128;  // bipartition columns (hA, hB, ps)
constexpr int KC   = 16;   // K per chunk
constexpr int ROWB = 80;   // staged row stride (bytes): hi[16] lo[16], then the row's
                           // FP32 scale for parts of even / odd ordinal at bytes 64 / 68
constexpr int BOX  = 32768;
constexpr int PF   = 2;    // chunks in flight per thread (plus one in registers)

struct __align__(16) MSmem {
  float         t3[BOX];
  unsigned char a[MM * ROWB];
  unsigned char b[NN * ROWB];
  double        eps[36];  // h1,h2,h3 (4 each) at 0,4,8; p4,p5,p6 (8 each) at 12,20,28
  double        red[8];
  int           sub[6], dim[6], off[6];
};

// bit position of each t3 index in the box layout (2 bits per hole, 3 per particle)
__device__ __forceinline__ int sh(int pos) { return pos < 3 ? 2 * pos : 6 + 3 * (pos - 3); }
__device__ __forceinline__ int ext(int pos) { return pos < 3 ? 4 : 8; }

__device__ __forceinline__ void pair_of(int s, int& a, int& b) {
  a = (s == 0) ? 1 : 0;
  b = (s == 2) ? 1 : 2;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for(int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ void fx_add(unsigned long long* acc, double p) {
  const bool         neg  = p < 0.0;
  const double       a    = fabs(p);
  unsigned long long hi   = (unsigned long long) a;
  const double       frac = a - (double) hi;
  unsigned long long lo   = (unsigned long long) (frac * 18446744073709551616.0);
  if(neg) {
    lo = ~lo + 1ull;
    hi = ~hi + (lo == 0ull ? 1ull : 0ull);
  }
  const unsigned long long old   = atomicAdd(&acc[0], lo);
  const unsigned long long carry = (old + lo < old) ? 1ull : 0ull;
  atomicAdd(&acc[1], hi + carry);
}

// One K part of a bipartition, as seen by one thread: its staged A row
// (m = tid) and half of B row (n = tid/2), their pre-split data and scales.
struct Part {
  const unsigned char* pa;  // A row data (null: zero row)
  const unsigned char* pb;  // B row data
  float                sa, sb;
  int                  nch;
};

struct Ring {
  uint4 a[4];  // 64 B: k0..7 [hi8 lo8], k8..15 [hi8 lo8]
  uint4 b[2];  // 32 B: this thread's 8-k half [hi8 lo8]
  float sa, sb;
};

// Part gp = 2*g + (d2 ? 1 : 0) of bipartition g for this thread.
__device__ Part part_setup(const MSmem& sm, const MixedDev& M, int gp, unsigned code) {
  const int g = gp >> 1, d2 = gp & 1;
  const int hs = g / 3, ps = g % 3;
  int       hA, hB, pA, pB;
  pair_of(hs, hA, hB);
  pair_of(ps, pA, pB);
  const int tid = threadIdx.x;
  const int m = tid, n = tid >> 1, half = tid & 1;
  const int gHS = 4 * sm.sub[hs] + (m & 3), gPA = 8 * sm.sub[3 + pA] + ((m >> 2) & 7),
            gPB = 8 * sm.sub[3 + pB] + (m >> 5);
  const int gHA = 4 * sm.sub[hA] + (n & 3), gHB = 4 * sm.sub[hB] + ((n >> 2) & 3),
            gPS = 8 * sm.sub[3 + ps] + (n >> 4);
  const bool mok = gHS < sm.dim[hs] && gPA < sm.dim[3 + pA] && gPB < sm.dim[3 + pB];
  const bool nok = gHA < sm.dim[hA] && gHB < sm.dim[hB] && gPS < sm.dim[3 + ps];
  const long O = M.O, V = M.V;
  const long HS = sm.off[hs] + (mok ? gHS : 0), PA = sm.off[3 + pA] + (mok ? gPA : 0),
             PB = sm.off[3 + pB] + (mok ? gPB : 0);
  const long HA = sm.off[hA] + (nok ? gHA : 0), HB = sm.off[hB] + (nok ? gHB : 0),
             PS = sm.off[3 + ps] + (nok ? gPS : 0);
  const float sg = (code & 2) ? -1.f : 1.f;
  // row index and operand array of c1*x - c2*y (both -> antisymmetrised array)
  auto choose = [](unsigned c, unsigned bx, unsigned by, int both, int single, long rx, long ry,
                   int& arr, long& row, float& coef) {
    const bool hx = (c & bx) != 0, hy = (c & by) != 0;
    arr  = hx && hy ? both : single;
    row  = hx ? rx : ry;
    coef = (hx || hy) ? (hx ? 1.f : -1.f) : 0.f;
  };
  int   aa, ab;
  long  ra, rb;
  float ca, cb;
  if(!d2) {
    choose(code, 4, 8, 1, 0, (PA * V + PB) * O + HS, (PB * V + PA) * O + HS, aa, ra, ca);
    choose(code, 16, 32, 3, 2, (HB * V + PS) * O + HA, (HA * V + PS) * O + HB, ab, rb, cb);
  }
  else {
    choose(code, 4, 8, 5, 4, (PB * V + PA) * O + HS, (PA * V + PB) * O + HS, aa, ra, ca);
    choose(code, 16, 32, 7, 6, (HA * O + HB) * V + PS, (HB * O + HA) * V + PS, ab, rb, cb);
  }
  const long K = d2 ? M.Kv : M.Ko;  // padded row length (elements)
  Part       P;
  P.nch = (int) (K / KC);
  P.pa  = (mok && ca != 0.f) ? M.v[aa] + ra * K * 4 : nullptr;
  P.pb  = (nok && cb != 0.f) ? M.v[ab] + rb * K * 4 + 32 * half : nullptr;
  P.sa  = P.pa ? sg * ca * ldexpf(1.f, -(int) M.e[aa][ra]) : 0.f;
  P.sb  = P.pb ? cb * ldexpf(1.f, -(int) M.e[ab][rb]) : 0.f;
  return P;
}

__device__ __forceinline__ void fetch(const Part& P, int ch, Ring& r) {
  const uint4 z = make_uint4(0, 0, 0, 0);
  if(P.pa) {
    const uint4* q = reinterpret_cast<const uint4*>(P.pa + (long) ch * 64);
    r.a[0] = q[0]; r.a[1] = q[1]; r.a[2] = q[2]; r.a[3] = q[3];
  }
  else {
    r.a[0] = z; r.a[1] = z; r.a[2] = z; r.a[3] = z;
  }
  if(P.pb) {
    const uint4* q = reinterpret_cast<const uint4*>(P.pb + (long) ch * 64);
    r.b[0] = q[0]; r.b[1] = q[1];
  }
  else {
    r.b[0] = z; r.b[1] = z;
  }
  r.sa = P.sa;
  r.sb = P.sb;
}

__device__ __forceinline__ float& row_scale(unsigned char* row, int parity) {
  return *reinterpret_cast<float*>(row + 64 + 4 * parity);
}

// Stages one chunk: A row m=tid as hi[k0..15] | lo[k0..15]; B row n=tid/2
// half.  The first chunk of a part also stores the row scales of that part in
// the slot of its ordinal's parity.
__device__ __forceinline__ void stage(MSmem& sm, const Ring& r, bool first, int parity) {
  const int tid = threadIdx.x;
  uint4*    ar  = reinterpret_cast<uint4*>(sm.a + tid * ROWB);
  ar[0] = r.a[0];  // hi k0..7
  ar[2] = r.a[1];  // lo k0..7
  ar[1] = r.a[2];  // hi k8..15
  ar[3] = r.a[3];  // lo k8..15
  const int n = tid >> 1, half = tid & 1;
  uint4*    br = reinterpret_cast<uint4*>(sm.b + n * ROWB);
  br[half]     = r.b[0];
  br[2 + half] = r.b[1];
  if(first) {
    row_scale(sm.a + tid * ROWB, parity) = r.sa;
    if(!half) row_scale(sm.b + n * ROWB, parity) = r.sb;
  }
}

// MFMA operand fragments of one staged chunk for this wave: rows 64w..64w+63
// (2 row tiles of 32) and all 4 column tiles, hi and lo halves.
struct Frags {
  h8 ahi[2], alo[2], bhi[4], blo[4];
};

__device__ __forceinline__ void read_frags(const MSmem& sm, Frags& F) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 31, kh = (lane >> 5) * 16;
#pragma unroll
  for(int i = 0; i < 2; i++) {
    const unsigned char* p = sm.a + (64 * wave + 32 * i + r) * ROWB + kh;
    F.ahi[i]               = *reinterpret_cast<const h8*>(p);
    F.alo[i]               = *reinterpret_cast<const h8*>(p + 32);
  }
#pragma unroll
  for(int j = 0; j < 4; j++) {
    const unsigned char* p = sm.b + (32 * j + r) * ROWB + kh;
    F.bhi[j]               = *reinterpret_cast<const h8*>(p);
    F.blo[j]               = *reinterpret_cast<const h8*>(p + 32);
  }
}

// 256x128x16 step on registers: three passes over the 8 accumulators so
// consecutive MFMAs are independent.
__device__ __forceinline__ void mfma_frags(const Frags& F, f16v (&acc)[2][4]) {
#pragma unroll
  for(int j = 0; j < 4; j++)
#pragma unroll
    for(int i = 0; i < 2; i++)
      acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(F.ahi[i], F.bhi[j], acc[i][j], 0, 0, 0);
#pragma unroll
  for(int j = 0; j < 4; j++)
#pragma unroll
    for(int i = 0; i < 2; i++)
      acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(F.ahi[i], F.blo[j], acc[i][j], 0, 0, 0);
#pragma unroll
  for(int j = 0; j < 4; j++)
#pragma unroll
    for(int i = 0; i < 2; i++)
      acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(F.alo[i], F.bhi[j], acc[i][j], 0, 0, 0);
}

// Adds the scaled accumulators of bipartition g into the box and clears them.
// D layout of v_mfma_f32_32x32x16_f16: register r of lane l holds row
// 8*(r/4) + 4*(l/32) + r%4, column l%32.
__device__ __forceinline__ void flush(MSmem& sm, int g, int parity, f16v (&acc)[2][4]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int hs = g / 3, ps = g % 3;
  int       hA, hB, pA, pB;
  pair_of(hs, hA, hB);
  pair_of(ps, pA, pB);
#pragma unroll
  for(int i = 0; i < 2; i++)
#pragma unroll
    for(int j = 0; j < 4; j++) {
      const int   n  = 32 * j + (lane & 31);
      const float sb = row_scale(sm.b + n * ROWB, parity);
      const int   en = ((n & 3) << sh(hA)) | (((n >> 2) & 3) << sh(hB)) | ((n >> 4) << sh(3 + ps));
#pragma unroll
      for(int q = 0; q < 16; q++) {
        const int m  = 64 * wave + 32 * i + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
        const int em = ((m & 3) << sh(hs)) | (((m >> 2) & 7) << sh(3 + pA)) | ((m >> 5) << sh(3 + pB));
        sm.t3[em | en] += acc[i][j][q] * row_scale(sm.a + m * ROWB, parity) * sb;
        acc[i][j][q] = 0.f;
      }
    }
}

__device__ __forceinline__ bool part_on(uint64_t d1, uint64_t d2, int gp) {
  const uint64_t c = (gp & 1) ? d2 : d1;
  return (c >> (6 * (gp >> 1))) & 1;
}
__device__ __forceinline__ unsigned part_code(uint64_t d1, uint64_t d2, int gp) {
  const uint64_t c = (gp & 1) ? d2 : d1;
  return (unsigned) (c >> (6 * (gp >> 1))) & 63u;
}

__device__ void box_contract(MSmem& sm, const MixedDev& M, uint64_t d1, uint64_t d2) {
  f16v acc[2][4];
#pragma unroll
  for(int i = 0; i < 2; i++)
#pragma unroll
    for(int j = 0; j < 4; j++)
#pragma unroll
      for(int q = 0; q < 16; q++) acc[i][j][q] = 0.f;
  // total chunks and first enabled part
  int total = 0, gp0 = -1;
  for(int gp = 0; gp < 18; gp++)
    if(part_on(d1, d2, gp)) {
      total += (gp & 1) ? M.Kv / KC : M.Ko / KC;
      if(gp0 < 0) gp0 = gp;
    }
  if(total == 0) return;
  // fetch cursor
  int  fgp = gp0, fch = 0;
  Part fp  = part_setup(sm, M, fgp, part_code(d1, d2, fgp));
  auto advance_fetch = [&]() {
    if(++fch == fp.nch) {
      fch = 0;
      do {
        ++fgp;
      } while(fgp < 18 && !part_on(d1, d2, fgp));
      if(fgp < 18) fp = part_setup(sm, M, fgp, part_code(d1, d2, fgp));
    }
  };
  Ring ring[PF];
#pragma unroll
  for(int s = 0; s < PF; s++)
    if(s < total) {
      fetch(fp, fch, ring[s]);
      advance_fetch();
    }
  // compute cursor.  Software pipeline per chunk c (fragments of c already in
  // registers F): stage c+1 into LDS, barrier, then issue the MFMAs of c in the
  // same block as the LDS reads of c+1, barrier.  The matrix pipe works through
  // the reads; only the write and the two barriers separate consecutive chunks.
  int   cgp = gp0, cch = 0, cn = (cgp & 1) ? M.Kv / KC : M.Ko / KC, cpart = 0;
  Frags F[2];
  __syncthreads();
  stage(sm, ring[0], true, 0);
  __syncthreads();
  read_frags(sm, F[0]);
  if(PF < total) {
    fetch(fp, fch, ring[0]);
    advance_fetch();
  }
  __syncthreads();
#pragma unroll 1
  for(int it0 = 0; it0 < total; it0 += 2 * PF) {
#pragma unroll
    for(int s = 0; s < 2 * PF; s++) {
      const int it = it0 + s;  // chunk in F[s & 1]; chunk it+1 is in ring[(s+1) % PF]
      if(it < total) {
        const bool more = it + 1 < total;
        if(more) {
          const bool nf = cch + 1 == cn;  // chunk it+1 starts the next part
          stage(sm, ring[(s + 1) % PF], nf, (cpart + (nf ? 1 : 0)) & 1);
          if(it + 1 + PF < total) {
            fetch(fp, fch, ring[(s + 1) % PF]);
            advance_fetch();
          }
        }
        __syncthreads();
        if(more) read_frags(sm, F[(s + 1) & 1]);
        mfma_frags(F[s & 1], acc);
        __syncthreads();
        if(++cch == cn) {
          flush(sm, cgp >> 1, cpart & 1, acc);
          // a one-chunk next part lets a fast wave stage part cpart+2 (same
          // scale slot) in its next iteration: all flush reads finish first
          __syncthreads();
          cch = 0;
          ++cpart;
          do {
            ++cgp;
          } while(cgp < 18 && !part_on(d1, d2, cgp));
          cn = (cgp & 1) ? M.Kv / KC : M.Ko / KC;
        }
      }
    }
  }
}

// Energy of the box.  E4 = sum d^2/D and E5 = E4 + sum d*s/D, with s the s1
// term; s is linear, so its 9 slots are processed in two passes (slots 0-5,
// 6-8) whose T1 and <hk hj||pc pb> blocks are staged (FP32) in the free
// operand staging area.  Slots are compile-time in each pass, so all index
// arithmetic folds.  Denominators and sums are FP64.
__device__ __forceinline__ double box_denom(const MSmem& sm, const int* idx) {
  return sm.eps[8 + idx[2]] + sm.eps[4 + idx[1]] + sm.eps[idx[0]] - sm.eps[28 + idx[5]] -
         sm.eps[20 + idx[4]] - sm.eps[12 + idx[3]];
}

template<int S0, int NS>
__device__ __forceinline__ double s1_pass(MSmem& sm, const MixedDev& M, uint64_t s1code,
                                          const int* lim) {
  static_assert(NS * (32 + 1024) * 4 <= MM * ROWB + NN * ROWB, "s1 staging too large");
  const int  tid = threadIdx.x;
  const long O = M.O, V = M.V;
  float*     T1s = reinterpret_cast<float*>(sm.a);  // [NS][8 pa][4 hi]
  float*     V1s = T1s + NS * 32;                   // [NS][1024] hk + 4 hj + 16 pc + 128 pb
  __syncthreads();                                  // staging area free
#pragma unroll 1
  for(int q = 0; q < NS; q++) {
    const int      slot = S0 + q;
    const unsigned c    = (unsigned) (s1code >> (4 * slot)) & 15u;
    const int      pa = slot / 3, hi = slot % 3;
    int            hj, hk, pb, pc;
    pair_of(hi, hj, hk);
    pair_of(pa, pb, pc);
    if(tid < 32) {
      const int gp = 8 * sm.sub[3 + pa] + (tid >> 2), gh = 4 * sm.sub[hi] + (tid & 3);
      float     v  = 0.f;
      if((c & 1) && gp < sm.dim[3 + pa] && gh < sm.dim[hi]) {
        const double t = M.t1[(long) (sm.off[3 + pa] + gp) * O + sm.off[hi] + gh];
        v              = (float) ((c & 2) ? -t : t);
      }
      T1s[q * 32 + tid] = v;
    }
    double x1[4], x2[4];
#pragma unroll
    for(int j = 0; j < 4; j++) {
      const int w = tid + NT * j;
      const int gk = 4 * sm.sub[hk] + (w & 3), gj = 4 * sm.sub[hj] + ((w >> 2) & 3),
                gc = 8 * sm.sub[3 + pc] + ((w >> 4) & 7), gb = 8 * sm.sub[3 + pb] + (w >> 7);
      x1[j] = 0.0;
      x2[j] = 0.0;
      if((c & 1) && gk < sm.dim[hk] && gj < sm.dim[hj] && gc < sm.dim[3 + pc] &&
         gb < sm.dim[3 + pb]) {
        const long K = sm.off[hk] + gk, J = sm.off[hj] + gj, PC = sm.off[3 + pc] + gc,
                   PB = sm.off[3 + pb] + gb;
        if(c & 4) x1[j] = M.vvoo[((PC * V + PB) * O + K) * O + J];
        if(c & 8) x2[j] = -M.vvoo[((PB * V + PC) * O + K) * O + J];
      }
    }
#pragma unroll
    for(int j = 0; j < 4; j++) V1s[q * 1024 + tid + NT * j] = (float) (x1[j] + x2[j]);
  }
  __syncthreads();
  double es = 0.0;
#pragma unroll 1
  for(int rr = 0; rr < BOX / NT; rr++) {
    const int e = tid + NT * rr;
    int       idx[6];
    bool      ok = true;
#pragma unroll
    for(int p = 0; p < 6; p++) {
      idx[p] = (e >> sh(p)) & (ext(p) - 1);
      ok     = ok && idx[p] < lim[p];
    }
    if(!ok) continue;
    float sv = 0.f;
#pragma unroll
    for(int q = 0; q < NS; q++) {
      const int slot = S0 + q;
      const int pa = slot / 3, hi = slot % 3;
      int       hj, hk, pb, pc;
      pair_of(hi, hj, hk);
      pair_of(pa, pb, pc);
      sv += T1s[q * 32 + idx[3 + pa] * 4 + idx[hi]] *
            V1s[q * 1024 + idx[hk] + 4 * idx[hj] + 16 * idx[3 + pc] + 128 * idx[3 + pb]];
    }
    es += (double) sm.t3[e] * (double) sv / box_denom(sm, idx);
  }
  return es;
}

__device__ void box_energy(MSmem& sm, const MixedDev& M, uint64_t s1code, double& e4,
                           double& e5) {
  const int tid = threadIdx.x;
  int       lim[6];
#pragma unroll
  for(int p = 0; p < 6; p++) lim[p] = sm.dim[p] - ext(p) * sm.sub[p];
  e4 = 0.0;
#pragma unroll 2
  for(int rr = 0; rr < BOX / NT; rr++) {
    const int e = tid + NT * rr;
    int       idx[6];
    bool      ok = true;
#pragma unroll
    for(int p = 0; p < 6; p++) {
      idx[p] = (e >> sh(p)) & (ext(p) - 1);
      ok     = ok && idx[p] < lim[p];
    }
    if(!ok) continue;
    const double d = (double) sm.t3[e];
    e4 += d * d / box_denom(sm, idx);
  }
  double es = 0.0;
  if(s1code & 0x111111ull) es += s1_pass<0, 6>(sm, M, s1code, lim);
  if(s1code & 0x111000000ull) es += s1_pass<6, 3>(sm, M, s1code, lim);
  e5                  = e4 + es;
  const int lane = tid & 63, wave = tid >> 6;
  e4 = wave_sum(e4);
  e5 = wave_sum(e5);
  if(lane == 0) {
    sm.red[wave]     = e4;
    sm.red[4 + wave] = e5;
  }
  __syncthreads();
  if(tid == 0) {
    e4 = ((sm.red[0] + sm.red[1]) + sm.red[2]) + sm.red[3];
    e5 = ((sm.red[4] + sm.red[5]) + sm.red[6]) + sm.red[7];
  }
}

__global__ void __launch_bounds__(NT) k_spatial_mixed(MixedDev M, RunDev R) {
  __shared__ MSmem     sm;
  __shared__ int       s_task;
  __shared__ long long s_local;
  const int            tid  = threadIdx.x;
  const int            home = blockIdx.x % kQueues;
#pragma unroll 1
  for(int qi = 0; qi < kQueues; qi++) {
    const int q = (home + qi) % kQueues;
#pragma unroll 1
    for(;;) {
      __syncthreads();
      if(tid == 0) {
        const unsigned long long t = atomicAdd(&R.qctr[q], 1ull);
        if((long long) t >= R.qtotal[q]) s_task = -1;
        else {
          int lo = R.qbeg[q], hi = R.qbeg[q + 1] - 1;
          while(lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if(R.qtick[mid] <= (long long) t) lo = mid;
            else hi = mid - 1;
          }
          s_task  = R.qtask[lo];
          s_local = (long long) t - R.qtick[lo];
        }
      }
      __syncthreads();
      const int ti = s_task;
      if(ti < 0) break;
      const TaskDev& T   = R.tasks[ti];
      long long      box = s_local >> 6;
      const int      inb = (int) (s_local & 63);
      bool           ok  = true;
      int            sub[6];
#pragma unroll
      for(int p = 0; p < 6; p++) {
        const int nb = T.nboxm[p];
        sub[p]       = 2 * (int) (box % nb) + ((inb >> p) & 1);
        box /= nb;
        ok = ok && sub[p] < T.nsubm[p];
      }
      if(!ok) continue;
#pragma unroll
      for(int p = 0; p < 6; p++)
        if(tid == p) {
          sm.sub[p] = sub[p];
          sm.dim[p] = T.dim[p];
          sm.off[p] = T.off[p];
        }
      for(int e = tid; e < BOX; e += NT) sm.t3[e] = 0.f;
      __syncthreads();
      if(tid < 36) {
        const int p = tid < 12 ? tid / 4 : 3 + (tid - 12) / 8;
        const int i = tid < 12 ? tid % 4 : (tid - 12) % 8;
        const int g = ext(p) * sm.sub[p] + i;
        sm.eps[tid] = g < sm.dim[p] ? (p < 3 ? M.eo : M.ev)[sm.off[p] + g] : 0.0;
      }
      const unsigned long long t0 = wall_clock64();
      box_contract(sm, M, T.d1code, T.d2code);
      __syncthreads();
      double e4, e5;
      box_energy(sm, M, T.s1code, e4, e5);
      if(tid == 0) {
        fx_add(&R.acc[4 * (size_t) ti], e4);
        fx_add(&R.acc[4 * (size_t) ti + 2], e5);
        if(R.cost) atomicAdd(&R.cost[ti], wall_clock64() - t0);
      }
    }
  }
}

// Per-row exponent: one wave per row, e = 13 - floor(log2(max|x|)).
__global__ void k_row_exp(const double* __restrict__ x, long rows, long Kin, short* __restrict__ e) {
  const long row  = (long) blockIdx.x * 4 + (threadIdx.x >> 6);
  const int  lane = threadIdx.x & 63;
  if(row >= rows) return;
  double m = 0.0;
  for(long k = lane; k < Kin; k += 64) m = fmax(m, fabs(x[row * Kin + k]));
#pragma unroll
  for(int o = 32; o >= 1; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if(lane == 0) {
    int ex = 0;
    if(m > 0.0) ex = 13 - ilogb(m);
    ex     = ex > 200 ? 200 : (ex < -200 ? -200 : ex);
    e[row] = (short) ex;
  }
}

// Split: y = x * 2^e[row] = hi + lo (FP16), written per 8 k as [hi8][lo8].
__global__ void k_split(const double* __restrict__ x, long rows, long Kin, long Kout,
                        const short* __restrict__ e, _Float16* __restrict__ out) {
  const long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= rows * Kout) return;
  const long   row = i / Kout, k = i % Kout;
  const double y   = k < Kin ? ldexp(x[row * Kin + k], (int) e[row]) : 0.0;
  const _Float16 h = (_Float16) y;
  const _Float16 l = (_Float16) (y - (double) h);
  const long   base = row * Kout * 2 + (k / 8) * 16 + (k % 8);
  out[base]     = h;
  out[base + 8] = l;
}

}  // namespace

int build_split(const SpatialBufs& B, int O, int V, SplitBufs& S, void* stream) {
  hipStream_t   st  = (hipStream_t) stream;
  const long    o = O, v = V, op = (O + 7) / 8 * 8, vp = (V + 7) / 8 * 8;
  const long    ko = (O + 15) / 16 * 16, kv = (V + 15) / 16 * 16;
  const double* src[kOperands] = {B.t2h, B.t2ha, B.ovk, B.ovka, B.xv, B.xva, B.t2p, B.t2pa};
  const long    rows[kOperands] = {v * v * o, v * v * o, o * v * o, o * v * o,
                                   v * v * o, v * v * o, o * o * v, o * o * v};
  for(int j = 0; j < kOperands; j++) {
    const long kin = j < 4 ? op : vp, kout = j < 4 ? ko : kv;
    S.rows[j]      = rows[j];
    hipError_t err = hipMalloc((void**) &S.e[j], sizeof(short) * rows[j]);
    if(err) return (int) err;
    err = hipMalloc((void**) &S.v[j], 4 * rows[j] * kout);
    if(err) return (int) err;
    hipLaunchKernelGGL(k_row_exp, dim3((unsigned) ((rows[j] + 3) / 4)), dim3(256), 0, st, src[j],
                       rows[j], kin, S.e[j]);
    const long n = rows[j] * kout;
    hipLaunchKernelGGL(k_split, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, st, src[j],
                       rows[j], kin, kout, S.e[j], reinterpret_cast<_Float16*>(S.v[j]));
    if((err = hipGetLastError())) return (int) err;
  }
  return (int) hipStreamSynchronize(st);
}

int launch_spatial_mixed(const MixedDev& M, const RunDev& R, int grid, void* stream) {
  hipLaunchKernelGGL(k_spatial_mixed, dim3(grid), dim3(NT), 0, (hipStream_t) stream, M, R);
  return (int) hipGetLastError();
}

}  // namespace xt
