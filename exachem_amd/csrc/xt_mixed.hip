// Mixed-precision (T) kernel for gfx950.
//
// Same task / bipartition decomposition as the FP64 kernel, with larger
// workgroup boxes and 16-bit matrix cores:
//   * a workgroup owns a (4,4,4 | 8,8,8) box of t3[h1,h2,h3,p4,p5,p6]
//     (32768 elements, FP32 in LDS, never in HBM);
//   * each bipartition g (single hole hs + particle pair | hole pair +
//     single particle ps) is a 256 x 128 GEMM: M = (hs,pA,pB) 4*8*8,
//     N = (hA,hB,ps) 4*4*8, K = h7 (d1) then p7 (d2);
//   * operands are pre-split FP16 pairs y = hi + lo of x * 2^e, with one
//     power-of-two exponent per (tile, tile, tile) block of rows (per-tile
//     amplitude scaling: the block's max |y| lies in [2^13, 2^14)); every
//     product is hi*hi + hi*lo + lo*hi on v_mfma_f32_32x32x16_f16 with FP32
//     accumulation (~22-bit products).  All rows of one contraction part of a
//     task lie in one block per operand, so the part's scale (both exponents
//     and the spin-orbital signs) is one number, applied when the part ends and
//     its result is added into the FP32 box;
//   * eight waves (two per SIMD, so one wave's LDS and VALU work overlaps the
//     other's MFMAs); wave w owns the 32 M rows of particle pB = w, loaded
//     from L2/HBM straight into registers in the MFMA operand layout.  The B
//     panel (shared by all waves) goes through a double-buffered LDS stage.
//     A register ring keeps PF chunks of loads in flight across part and
//     bipartition boundaries; every thread issues the same loads for every
//     chunk (rows outside the tile read a zero row), so the compiler's wait
//     counts follow the ring; one barrier per K chunk;
//   * MFMAs are not skipped: rows and columns outside the task tile read the
//     zero row and add zeros (a branch per MFMA costs more issue time than
//     the skipped work; a branch per chunk for a wave whose row tile is
//     outside makes the compiler spill 25 registers);
//   * the flush into the box adds precomputed per-lane, per-row and
//     per-column index parts (disjoint bit fields of the box index); the d1
//     part of a bipartition carries its accumulators into the d2 part, so a
//     bipartition flushes once;
//   * energy: one FP64 reciprocal of the denominator per element (from
//     separable hole and particle sums), E4 in FP64, then t3/D is kept in
//     place for the s1 sweeps (FP32 products, FP64 sums); task partials are
//     accumulated exactly like the FP64 kernel.
#include <hip/hip_runtime.h>

#include "xt_device.h"
#include "xt_fixed.h"
#include "xt_queue.h"
#include "xt_tasksums.h"

namespace xt {
namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float    f16v __attribute__((ext_vector_type(16)));

constexpr int NT  = 512;
constexpr int NW  = NT / 64;  // waves; wave w owns M rows 32w..32w+31
constexpr int NN  = 128;      // bipartition columns (hA, hB, ps)
constexpr int KC  = 16;       // K per chunk (one MFMA K step)
constexpr int BOX = 32768;
constexpr int PF  = 4;        // ring depth: chunks whose loads are in flight per thread
constexpr int S1P = 3;        // s1 slots staged per sweep
// the unrolled ring loop picks stage and fragment buffers by s & 1, which is
// the chunk's parity only when the ring depth is even
static_assert(PF % 2 == 0, "ring depth must be even");
// chunk it is multiplied from ring[it % PF] while chunk it+1's scale and
// chunk it+2's B piece are read from the next two slots
static_assert(PF >= 3, "the ring must hold the current chunk and the next two");

struct __align__(16) MSmem {
  float t3[BOX];
  union {
    struct {
      uint4 hi[2][2][NN];  // [buffer][k half][column]: FP16 hi of 8 k
      uint4 lo[2][2][NN];  // FP16 lo of the same 8 k
    } g;
    struct {
      double dh[64];   // e_h1 + e_h2 + e_h3 over the box's holes   (index h1 | h2<<2 | h3<<4)
      double dp[512];  // e_p4 + e_p5 + e_p6 over the box's particles (p4 | p5<<3 | p6<<6)
    } d;
    struct {
      float t1[S1P][32];    // [slot][pa*4 + hi]
      float v1[S1P][1088];  // [slot][hk + 4 hj + 16 pc + 128 pb]; p4 slots (hk + 4 hj) * 68 + 8 pc + pb
    } s;
  } u;
  double eps[36];  // h1,h2,h3 (4 each) at 0,4,8; p4,p5,p6 (8 each) at 12,20,28
  double red[2 * NW];
  float  scale[18];  // scale of each contraction part of the box's task
  int    sub[6], dim[6], off[6], tile[6];
};

// bit position of each t3 index in the box layout (2 bits per hole, 3 per particle)
__host__ __device__ constexpr int sh(int pos) { return pos < 3 ? 2 * pos : 6 + 3 * (pos - 3); }
__host__ __device__ constexpr int ext(int pos) { return pos < 3 ? 4 : 8; }
__host__ __device__ constexpr int pair_a(int s) { return s == 0 ? 1 : 0; }
__host__ __device__ constexpr int pair_b(int s) { return s == 2 ? 1 : 2; }

__device__ __forceinline__ void pair_of(int s, int& a, int& b) {
  a = pair_a(s);
  b = pair_b(s);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for(int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ h8 as_h8(uint4 u) { return __builtin_bit_cast(h8, u); }

// One K part of a bipartition as seen by one thread: the A row of its lane
// (M row 32w + lane%32, k half lane/32), its 16-byte piece of a B column
// (column tid/4, piece tid%4 = [hi k0-7 | lo k0-7 | hi k8-15 | lo k8-15]) and
// the part's scale (the same for every thread).  Rows outside the tile read
// the zero row, so every thread issues the same loads for every chunk and the
// compiler's wait counts follow the ring.
struct Part {
  const unsigned char* pa;
  const unsigned char* pb;
  float                sc;  // sign * 2^-(eA + eB)
  int                  nch;
};

// One ring slot: the loads of one K chunk plus what its consumers need.
struct Slot {
  uint4 a[2];  // hi, lo of this lane's 8 k
  uint4 b;     // this thread's B piece
  float sc;    // scale of the chunk's part
};

// Operand array, row and coefficient of the term c1*x - c2*y of a
// spin-orbital element (bits bx, by of the part code): both terms -> the
// antisymmetrised array (holding x - y), one term -> the plain array.
__host__ __device__ __forceinline__ void choose_term(unsigned c, unsigned bx, unsigned by, int both,
                                                     int single, int rx, int ry, int& arr, int& row,
                                                     float& coef) {
  const bool hx = (c & bx) != 0, hy = (c & by) != 0;
  arr  = hx && hy ? both : single;
  row  = hx ? rx : ry;
  coef = (hx || hy) ? (hx ? 1.f : -1.f) : 0.f;
}

// Part gp = 2*g + (d2 ? 1 : 0) of bipartition g for this thread; gp = 18 is
// the zero part that pads the chunk count to a multiple of the ring depth.
__device__ __forceinline__ Part part_setup(const MSmem& sm, const MixedDev& M, int gp,
                                           unsigned code) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int n = tid >> 2, piece = tid & 3;
  Part      P;
  if(gp >= 18) {
    P.pa  = M.zero + 32 * (lane >> 5);
    P.pb  = M.zero + 16 * piece;
    P.sc  = 0.f;
    P.nch = 1;
    return P;
  }
  const int g = gp >> 1, d2 = gp & 1;
  const int hs = g / 3, ps = g % 3;
  int       hA, hB, pA, pB;
  pair_of(hs, hA, hB);
  pair_of(ps, pA, pB);
  const int  O = M.O, V = M.V;      // row indices stay below 2^31 (V*V*O rows at most)
  const long K = d2 ? M.Kv : M.Ko;  // padded row length (elements)
  // operand array, row, block and coefficient of c1*x - c2*y (both terms ->
  // the antisymmetrised array, which holds x - y)
  P.nch = (int) (K / KC);
  // the box's tile data of the six positions, read together (one LDS wait);
  // the row tests below use & so that no read sits behind a branch
  const int sHS = sm.sub[hs], sPA = sm.sub[3 + pA], sPB = sm.sub[3 + pB];
  const int sHA = sm.sub[hA], sHB = sm.sub[hB], sPS = sm.sub[3 + ps];
  const int dHS = sm.dim[hs], dPA = sm.dim[3 + pA], dPB = sm.dim[3 + pB];
  const int dHA = sm.dim[hA], dHB = sm.dim[hB], dPS = sm.dim[3 + ps];
  const int oHS = sm.off[hs], oPA = sm.off[3 + pA], oPB = sm.off[3 + pB];
  const int oHA = sm.off[hA], oHB = sm.off[hB], oPS = sm.off[3 + ps];
  P.sc          = sm.scale[gp];  // per task, copied to LDS at the box start
  int   aa, ra, ab, rb;
  float ca, cb;
  {
    const int  m   = 32 * wave + (lane & 31);
    const int  gHS = 4 * sHS + (m & 3), gPA = 8 * sPA + ((m >> 2) & 7), gPB = 8 * sPB + (m >> 5);
    const bool mok = (gHS < dHS) & (gPA < dPA) & (gPB < dPB);
    const int  HS = oHS + (mok ? gHS : 0), PA = oPA + (mok ? gPA : 0), PB = oPB + (mok ? gPB : 0);
    if(!d2)  // t2h[a][b][i]: rows (pA,pB,hs) or (pB,pA,hs)
      choose_term(code, 4, 8, 1, 0, (PA * V + PB) * O + HS, (PB * V + PA) * O + HS, aa, ra, ca);
    else  // xv[c][b][i]: rows (pB,pA,hs) or (pA,pB,hs)
      choose_term(code, 4, 8, 5, 4, (PB * V + PA) * O + HS, (PA * V + PB) * O + HS, aa, ra, ca);
    P.pa = (mok & (ca != 0.f)) ? M.v[aa] + (long) ra * K * 4 + 32 * (lane >> 5)
                               : M.zero + 32 * (lane >> 5);
  }
  {
    const int  gHA = 4 * sHA + (n & 3), gHB = 4 * sHB + ((n >> 2) & 3), gPS = 8 * sPS + (n >> 4);
    const bool nok = (gHA < dHA) & (gHB < dHB) & (gPS < dPS);
    const int  HA = oHA + (nok ? gHA : 0), HB = oHB + (nok ? gHB : 0), PS = oPS + (nok ? gPS : 0);
    if(!d2)  // ovk[i][a][j]: rows (hB,ps,hA) or (hA,ps,hB)
      choose_term(code, 16, 32, 3, 2, (HB * V + PS) * O + HA, (HA * V + PS) * O + HB, ab, rb, cb);
    else  // t2p[i][j][a]: rows (hA,hB,ps) or (hB,hA,ps)
      choose_term(code, 16, 32, 7, 6, (HA * O + HB) * V + PS, (HB * O + HA) * V + PS, ab, rb, cb);
    P.pb = (nok & (cb != 0.f)) ? M.v[ab] + (long) rb * K * 4 + 16 * piece : M.zero + 16 * piece;
  }
  return P;
}

// Issues the loads of chunk ch of part P (always the same three loads).
__device__ __forceinline__ void fetch(const Part& P, int ch, Slot& s) {
  const uint4* q = reinterpret_cast<const uint4*>(P.pa + (long) ch * 64);
  s.a[0]         = q[0];
  s.a[1]         = q[1];
  s.b            = *reinterpret_cast<const uint4*>(P.pb + (long) ch * 64);
  s.sc           = P.sc;
}

// Writes a chunk's B pieces into LDS buffer `buf`.
__device__ __forceinline__ void stage(MSmem& sm, const Slot& s, int buf) {
  const int tid = threadIdx.x, n = tid >> 2, piece = tid & 3;
  uint4*    plane = (piece & 1) ? &sm.u.g.lo[buf][piece >> 1][n] : &sm.u.g.hi[buf][piece >> 1][n];
  *plane          = s.b;
}

struct Frags {
  h8 bhi[4], blo[4];
};

__device__ __forceinline__ void read_frags(const MSmem& sm, int buf, Frags& F) {
  const int lane = threadIdx.x & 63, kh = lane >> 5;
#pragma unroll
  for(int j = 0; j < 4; j++) {
    F.bhi[j] = as_h8(sm.u.g.hi[buf][kh][32 * j + (lane & 31)]);
    F.blo[j] = as_h8(sm.u.g.lo[buf][kh][32 * j + (lane & 31)]);
  }
}

// 32x128x16 of this wave: three passes over its 4 accumulators.  Column
// tiles outside the task tile are not skipped: their B rows are the zero row,
// and a branch per MFMA costs more issue time than the skipped MFMA.
__device__ __forceinline__ void mfma_chunk(const Slot& s, const Frags& F, f16v (&acc)[4]) {
  const h8 ahi = as_h8(s.a[0]), alo = as_h8(s.a[1]);
#pragma unroll
  for(int j = 0; j < 4; j++) acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, F.bhi[j], acc[j], 0, 0, 0);
#pragma unroll
  for(int j = 0; j < 4; j++) acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, F.blo[j], acc[j], 0, 0, 0);
#pragma unroll
  for(int j = 0; j < 4; j++) acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(alo, F.bhi[j], acc[j], 0, 0, 0);
}

// Adds the scaled accumulators of a bipartition with row hole HS into the
// box and clears them.  D layout of v_mfma_f32_32x32x16_f16: register q of
// lane l holds row 8*(q/4) + 4*(l/32) + q%4, column l%32.  With
//   m = 32 wave + 8 (q/4) + 4 (lane/32) + q%4: hs = m & 3, pA = (m >> 2) & 7, pB = m >> 5
//   n = 32 j + lane%32:                       hA = n & 3, hB = (n >> 2) & 3, ps = n >> 4
// the t3 index of (q, j) is a per-lane base plus a j part plus a q/4 part plus
// (q%4) << sh(HS).  For HS = 0 (h1, the fastest t3 index) the four q%4 are
// consecutive floats and go through one 16-byte LDS read and write.
template<int HS>
__device__ __forceinline__ void flush_h(MSmem& sm, int ps, float sc, f16v (&acc)[4]) {
  constexpr int hA = pair_a(HS), hB = pair_b(HS), shs = sh(HS);
  const int     pA = pair_a(ps), pB = pair_b(ps);
  const int     lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int     spa = sh(3 + pA), sps = sh(3 + ps);
  const int     base = ((lane >> 5) << spa) | (wave << sh(3 + pB)) | ((lane & 3) << sh(hA)) |
                   (((lane >> 2) & 3) << sh(hB)) | (((lane >> 4) & 1) << sps);
#pragma unroll
  for(int qq = 0; qq < 4; qq++)
#pragma unroll
    for(int j = 0; j < 4; j++) {
      const int o = base + ((2 * j) << sps) + ((2 * qq) << spa);
      if constexpr(HS == 0) {
        float4* p = reinterpret_cast<float4*>(&sm.t3[o]);  // o is a multiple of 4
        float4  v = *p;
        v.x += acc[j][4 * qq + 0] * sc;
        v.y += acc[j][4 * qq + 1] * sc;
        v.z += acc[j][4 * qq + 2] * sc;
        v.w += acc[j][4 * qq + 3] * sc;
        *p = v;
      }
      else {
#pragma unroll
        for(int r = 0; r < 4; r++) sm.t3[o + (r << shs)] += acc[j][4 * qq + r] * sc;
      }
    }
#pragma unroll
  for(int j = 0; j < 4; j++)
#pragma unroll
    for(int q = 0; q < 16; q++) acc[j][q] = 0.f;
}

__device__ __forceinline__ void flush(MSmem& sm, int g, float sc, f16v (&acc)[4]) {
  switch(g / 3) {
    case 0: flush_h<0>(sm, g % 3, sc, acc); break;
    case 1: flush_h<1>(sm, g % 3, sc, acc); break;
    default: flush_h<2>(sm, g % 3, sc, acc); break;
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

// Whether this wave's row tile (pB = wave) of bipartition g lies inside the
// task tile; a wave outside it has only zero rows and skips its flushes.
__device__ __forceinline__ unsigned tile_mask(const MSmem& sm, int g) {
  const int ps = g % 3, pB = pair_b(ps), wave = threadIdx.x >> 6;
  return 8 * sm.sub[3 + pB] + wave < sm.dim[3 + pB] ? 1u : 0u;
}

__device__ __forceinline__ void box_contract(MSmem& sm, const MixedDev& M, uint64_t d1, uint64_t d2) {
  f16v acc[4];
#pragma unroll
  for(int j = 0; j < 4; j++)
#pragma unroll
    for(int q = 0; q < 16; q++) acc[j][q] = 0.f;
  int total = 0, gp0 = -1;
  for(int gp = 0; gp < 18; gp++)
    if(part_on(d1, d2, gp)) {
      total += (gp & 1) ? M.Kv / KC : M.Ko / KC;
      if(gp0 < 0) gp0 = gp;
    }
  if(total == 0) return;
  // chunk count padded with zero chunks to a multiple of the ring depth, so
  // the loop body is the same for every chunk
  const int total_pad = (total + PF - 1) / PF * PF;
  // fetch cursor (runs PF chunks ahead of the MFMAs); past the last part it
  // stays on the zero part
  int  fgp = gp0, fch = 0;
  Part fp  = part_setup(sm, M, fgp, part_code(d1, d2, fgp));
  auto fetch_next = [&](Slot& s) {
    fetch(fp, fch, s);
    if(++fch == fp.nch) {
      fch = 0;
      if(fgp < 18) {
        do {
          ++fgp;
        } while(fgp < 18 && !part_on(d1, d2, fgp));
        fp = part_setup(sm, M, fgp, fgp < 18 ? part_code(d1, d2, fgp) : 0u);
      }
    }
  };
  Slot ring[PF];
#pragma unroll
  for(int s = 0; s < PF; s++) fetch_next(ring[s]);
  // compute cursor
  int      cgp = gp0, cch = 0, cn = (cgp & 1) ? M.Kv / KC : M.Ko / KC;
  unsigned cmask = __builtin_amdgcn_readfirstlane(tile_mask(sm, cgp >> 1));
  Frags    F[2];
  // prologue: chunks 0 and 1 staged, fragments of chunk 0 read
  __syncthreads();  // previous box's energy sweeps are done with the union area
  stage(sm, ring[0], 0);
  stage(sm, ring[1], 1);
  __syncthreads();
  read_frags(sm, 0, F[0]);
  __syncthreads();
  // iteration it (chunk it: A in ring[it % PF], B fragments in F[it & 1]):
  //   read fragments of chunk it+1 (buffer (it+1)&1) while the MFMAs of
  //   chunk it run, stage chunk it+2 into buffer it&1 (its fragments were
  //   read before the previous barrier), flush at the end of a part, refill
  //   ring[it % PF] with chunk it+PF, barrier.
#pragma unroll 1
  for(int it0 = 0; it0 < total_pad; it0 += PF) {
#pragma unroll
    for(int s = 0; s < PF; s++) {
      read_frags(sm, (s + 1) & 1, F[(s + 1) & 1]);
      mfma_chunk(ring[s], F[s & 1], acc);
      stage(sm, ring[(s + 2) % PF], s & 1);
      if(++cch == cn) {
        int ngp = cgp;
        do {
          ++ngp;
        } while(ngp < 18 && !part_on(d1, d2, ngp));
        // A d1 part followed by the d2 part of the same bipartition adds into
        // the same box elements: its accumulators carry over, rescaled by the
        // ratio of the two part scales (both +-2^k, so the ratio is exact;
        // the next chunk, first of the d2 part, holds the d2 scale).  Ratios
        // outside 2^+-60 flush instead, keeping the FP32 range.
        bool carry = false;
        if(!(cgp & 1) && ngp == cgp + 1) {
          const float r = ring[s].sc / ring[(s + 1) % PF].sc;
          carry         = fabsf(r) >= 0x1p-60f && fabsf(r) <= 0x1p60f;
          if(carry)
#pragma unroll
            for(int j = 0; j < 4; j++)
#pragma unroll
              for(int q = 0; q < 16; q++) acc[j][q] *= r;
        }
        if(!carry && cmask) flush(sm, cgp >> 1, ring[s].sc, acc);
        cch = 0;
        cgp = ngp;
        cn    = cgp < 18 ? ((cgp & 1) ? M.Kv / KC : M.Ko / KC) : 0x7fffffff;
        cmask = __builtin_amdgcn_readfirstlane(cgp < 18 ? tile_mask(sm, cgp >> 1) : 0u);
      }
      fetch_next(ring[s]);
      __syncthreads();
    }
  }
}

// Denominator reciprocal in FP64: hardware estimate refined by two Newton
// steps (error below one ulp; |D| is bounded away from zero by the gap).
__device__ __forceinline__ double recip(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r        = fma(r, e, r);
  e        = fma(-d, r, 1.0);
  return fma(r, e, r);
}

// s1 term of S1P slots starting at S0, staged in LDS: returns this thread's
// sum over its elements of (t3/D) * s.  W = t3/D is already in sm.t3 (zero
// outside the tile), T1 and V blocks are zero outside the tile.
//
// The thread's elements e = tid + NT*rr are every (p5, p6) = (rr & 7, rr >> 3)
// for its own (h1, h2, h3, p4) (the bits of tid), and the slots of one sweep
// share their particle position pa = S0/3.  So a sweep needs one T1 value and
// 64 V values per slot (pa = p4), or 8 and 8 (pa = p5, p6): each is read from
// LDS once instead of once per element.  Per element the slots are summed in
// FP32 in slot order and the products W*s added in FP64 in element order, as
// an element-by-element loop would.
template<int S0>
__device__ __forceinline__ double s1_sweep(const MSmem& sm) {
  static_assert(S0 % 3 == 0 && S1P == 3, "a sweep covers the three slots of one pa");
  constexpr int pa = S0 / 3;
  const int     tid = threadIdx.x;
  const int     hidx[3] = {tid & 3, (tid >> 2) & 3, (tid >> 4) & 3};
  const int     p4 = (tid >> 6) & 7;
  int           vb[S1P], th[S1P];  // per slot: h part of the V index, T1 hole index
#pragma unroll
  for(int q = 0; q < S1P; q++) {
    const int hi = (S0 + q) % 3;
    int       hj, hk;
    pair_of(hi, hj, hk);
    vb[q] = hidx[hk] + 4 * hidx[hj];
    th[q] = hidx[hi];
  }
  double es = 0.0;
  if constexpr(pa == 0) {  // T1 by p4 (one value per slot); V by (pc, pb) = (p6, p5)
    float t[S1P];
#pragma unroll
    for(int q = 0; q < S1P; q++) t[q] = sm.u.s.t1[q][p4 * 4 + th[q]];
    // these slots are stored with pb (= p5) contiguous: the 8 values of one p6
    // are two 16-byte reads
#pragma unroll 1
    for(int p6 = 0; p6 < 8; p6++) {
      float4 v[S1P][2];
#pragma unroll
      for(int q = 0; q < S1P; q++) {
        const float4* r = reinterpret_cast<const float4*>(&sm.u.s.v1[q][vb[q] * 68 + 8 * p6]);
        v[q][0]         = r[0];
        v[q][1]         = r[1];
      }
#pragma unroll
      for(int p5 = 0; p5 < 8; p5++) {
        float sv = 0.f;
#pragma unroll
        for(int q = 0; q < S1P; q++) {
          const float4 h = v[q][p5 >> 2];
          const int    i = p5 & 3;
          sv += t[q] * (i == 0 ? h.x : i == 1 ? h.y : i == 2 ? h.z : h.w);
        }
        es += (double) (sm.t3[tid + NT * (p5 + 8 * p6)] * sv);
      }
    }
  }
  else {
    // pa = p5: T1 by p5, V by (pc, pb) = (p6, p4); pa = p6: T1 by p6, V by (p5, p4).
    // The operand indexed by p5 is held for the whole sweep (constant indices
    // in the unrolled p5 loop); the one indexed by p6 is read per p6.
    float hold[S1P][8];
#pragma unroll
    for(int q = 0; q < S1P; q++)
#pragma unroll
      for(int k = 0; k < 8; k++)
        hold[q][k] = pa == 1 ? sm.u.s.t1[q][k * 4 + th[q]] : sm.u.s.v1[q][vb[q] + 16 * k + 128 * p4];
#pragma unroll 1
    for(int p6 = 0; p6 < 8; p6++) {
      float per[S1P];
#pragma unroll
      for(int q = 0; q < S1P; q++)
        per[q] = pa == 1 ? sm.u.s.v1[q][vb[q] + 16 * p6 + 128 * p4] : sm.u.s.t1[q][p6 * 4 + th[q]];
#pragma unroll
      for(int p5 = 0; p5 < 8; p5++) {
        float sv = 0.f;
#pragma unroll
        for(int q = 0; q < S1P; q++) sv += pa == 1 ? hold[q][p5] * per[q] : per[q] * hold[q][p5];
        es += (double) (sm.t3[tid + NT * (p5 + 8 * p6)] * sv);
      }
    }
  }
  return es;
}

// Loads (into registers) the T1 entries and <hk hj||pc pb> blocks of slots
// S0..S0+S1P-1.  Consecutive threads read consecutive hj, the contiguous index
// of vvoo[c][b][k][j].
constexpr int S1V = 1024 / NT;  // block elements per thread and slot
struct S1Regs {
  double v[S1P][S1V];
  double t[S1P];
};

template<int S0>
__device__ __forceinline__ void s1_load(const MSmem& sm, const MixedDev& M, uint64_t s1code,
                                        S1Regs& R) {
  const int  tid = threadIdx.x;
  const long O = M.O, V = M.V;
#pragma unroll
  for(int q = 0; q < S1P; q++) {
    const int slot = S0 + q;
    R.t[q]         = 0.0;
#pragma unroll
    for(int j = 0; j < S1V; j++) R.v[q][j] = 0.0;
    if(slot >= 9) continue;
    const unsigned c  = (unsigned) (s1code >> (4 * slot)) & 15u;
    const int      pa = slot / 3, hi = slot % 3;
    int            hj, hk, pb, pc;
    pair_of(hi, hj, hk);
    pair_of(pa, pb, pc);
    if(!(c & 1)) continue;
    if(tid < 32) {
      const int gp = 8 * sm.sub[3 + pa] + (tid >> 2), gh = 4 * sm.sub[hi] + (tid & 3);
      if(gp < sm.dim[3 + pa] && gh < sm.dim[hi]) {
        const double t = M.t1[(long) (sm.off[3 + pa] + gp) * O + sm.off[hi] + gh];
        R.t[q]         = (c & 2) ? -t : t;
      }
    }
#pragma unroll
    for(int j = 0; j < S1V; j++) {
      const int w  = tid + NT * j;
      const int gj = 4 * sm.sub[hj] + (w & 3), gk = 4 * sm.sub[hk] + ((w >> 2) & 3),
                gc = 8 * sm.sub[3 + pc] + ((w >> 4) & 7), gb = 8 * sm.sub[3 + pb] + (w >> 7);
      if(gk < sm.dim[hk] && gj < sm.dim[hj] && gc < sm.dim[3 + pc] && gb < sm.dim[3 + pb]) {
        const long K = sm.off[hk] + gk, J = sm.off[hj] + gj, PC = sm.off[3 + pc] + gc,
                   PB = sm.off[3 + pb] + gb;
        double x = 0.0;
        if(c & 4) x = M.vvoo[((PC * V + PB) * O + K) * O + J];
        if(c & 8) x -= M.vvoo[((PB * V + PC) * O + K) * O + J];
        R.v[q][j] = x;
      }
    }
  }
}

// Stores S1P slots starting at S0.  The slots of particle position p4 (S0 = 0)
// go to (k + 4 j) * 68 + 8 c + b, so that their sweep reads b (= p5) runs;
// 68 rather than 64 keeps the stores of a wave on distinct banks but for pairs.
template<int S0>
__device__ __forceinline__ void s1_store(MSmem& sm, const S1Regs& R) {
  const int tid = threadIdx.x;
#pragma unroll
  for(int q = 0; q < S1P; q++) {
    if(tid < 32) sm.u.s.t1[q][tid] = (float) R.t[q];
#pragma unroll
    for(int j = 0; j < S1V; j++) {
      const int w = tid + NT * j;  // (j, k, c, b) -> stored as k + 4 j + 16 c + 128 b
      const int o = S0 == 0 ? (((w >> 2) & 3) + 4 * (w & 3)) * 68 + 8 * ((w >> 4) & 7) + (w >> 7)
                            : ((w >> 2) & 3) + 4 * (w & 3) + (w & ~15);
      sm.u.s.v1[q][o] = (float) R.v[q][j];
    }
  }
}

// Energy of the box: E4 = sum t3^2/D and E5 = E4 + sum t3*s/D over the tile
// elements of the box (FP64 sums of per-element terms; t3 and s in FP32).
__device__ __forceinline__ void box_energy(MSmem& sm, const MixedDev& M, uint64_t s1code, double& e4,
                           double& e5) {
  const int tid = threadIdx.x;
  S1Regs    R;
  s1_load<0>(sm, M, s1code, R);  // in flight during the E4 sweep
  // separable denominators, in the order of the reference's sum
  for(int x = tid; x < 64 + 512; x += NT) {
    if(x < 64) sm.u.d.dh[x] = (sm.eps[8 + (x >> 4)] + sm.eps[4 + ((x >> 2) & 3)]) + sm.eps[x & 3];
    else {
      const int y    = x - 64;
      sm.u.d.dp[y] = (sm.eps[28 + (y >> 6)] + sm.eps[20 + ((y >> 3) & 7)]) + sm.eps[12 + (y & 7)];
    }
  }
  __syncthreads();
  int lim[6];
#pragma unroll
  for(int p = 0; p < 6; p++) lim[p] = sm.dim[p] - ext(p) * sm.sub[p];
  e4 = 0.0;
#pragma unroll 2
  for(int rr = 0; rr < BOX / NT; rr++) {
    const int e  = tid + NT * rr;
    bool      ok = true;
#pragma unroll
    for(int p = 0; p < 6; p++) ok = ok && ((e >> sh(p)) & (ext(p) - 1)) < lim[p];
    float w = 0.f;
    if(ok) {
      const double d   = (double) sm.t3[e];
      const double inv = recip(sm.u.d.dh[e & 63] - sm.u.d.dp[e >> 6]);
      const double di  = d * inv;
      e4 += d * di;
      w = (float) di;
    }
    sm.t3[e] = w;
  }
  double es = 0.0;
  if(s1code & 0x111111111ull) {
    __syncthreads();  // denominators no longer needed; W complete
    s1_store<0>(sm, R);
    s1_load<S1P>(sm, M, s1code, R);
    __syncthreads();
    es += s1_sweep<0>(sm);
    __syncthreads();
    s1_store<S1P>(sm, R);
    s1_load<2 * S1P>(sm, M, s1code, R);
    __syncthreads();
    es += s1_sweep<S1P>(sm);
    __syncthreads();
    s1_store<2 * S1P>(sm, R);
    __syncthreads();
    es += s1_sweep<2 * S1P>(sm);
  }
  e5                  = e4 + es;
  const int lane = tid & 63, wave = tid >> 6;
  e4 = wave_sum(e4);
  e5 = wave_sum(e5);
  if(lane == 0) {
    sm.red[wave]      = e4;
    sm.red[NW + wave] = e5;
  }
  __syncthreads();
  if(tid == 0) {
    double a = 0.0, b = 0.0;
#pragma unroll
    for(int w = 0; w < NW; w++) {
      a += sm.red[w];
      b += sm.red[NW + w];
    }
    e4 = a;
    e5 = b;
  }
}

// Persistent kernel: workgroup b serves queue b % kQueues first (dispatch is
// round-robin over the 8 XCDs, so that is its own XCD's queue), then steals
// from the others; one ticket = one box.  Thread 0 claims and decodes each
// ticket (its entry searched forward from its previous one) and keeps the
// running exact sums of the task its boxes belong to, which go to the task's
// global accumulator when the task changes and when the workgroup leaves.
__global__ void __launch_bounds__(NT) k_spatial_mixed(MixedDev M, RunDev R) {
  __shared__ MSmem              sm;
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
      __syncthreads();
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
          // ticket -> box: the low bits select the box inside its group
          int sub[6];
          if(ticket_box((long long) t - R.qtick[e], T.nsubm, R.glog, sub)) {
#pragma unroll
            for(int p = 0; p < 6; p++) {
              sm.sub[p]  = sub[p];
              sm.dim[p]  = T.dim[p];
              sm.off[p]  = T.off[p];
              sm.tile[p] = T.tile[p];
            }
          }
          else ti = -2;  // a ticket past the task's box grid (its last group overhangs)
        }
        s_task = ti;
      }
      __syncthreads();
      const int ti = s_task;
      if(ti == -1) break;
      if(ti == -2) continue;
      if(tid >= 64 && tid < 64 + 18) sm.scale[tid - 64] = M.pscale[18 * (size_t) ti + tid - 64];
      {
        float4* t4 = reinterpret_cast<float4*>(sm.t3);
        for(int e = tid; e < BOX / 4; e += NT) t4[e] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
      if(tid < 36) {
        const int p = tid < 12 ? tid / 4 : 3 + (tid - 12) / 8;
        const int i = tid < 12 ? tid % 4 : (tid - 12) % 8;
        const int g = ext(p) * sm.sub[p] + i;
        sm.eps[tid] = g < sm.dim[p] ? (p < 3 ? M.eo : M.ev)[sm.off[p] + g] : 0.0;
      }
      __syncthreads();
      const TaskDev&           T  = R.tasks[ti];
      const unsigned long long t0 = wall_clock64();
      box_contract(sm, M, T.d1code, T.d2code);
      __syncthreads();
      double e4, e5;
      box_energy(sm, M, T.s1code, e4, e5);
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

// Per-row exponent: one wave per row, e = 13 - floor(log2(max|x|)); an
// all-zero row puts no constraint on its block (kNoRow).
constexpr int kNoRow = 200;

__global__ void k_row_exp(const double* __restrict__ x, long rows, long Kin, short* __restrict__ e) {
  const long row  = (long) blockIdx.x * 4 + (threadIdx.x >> 6);
  const int  lane = threadIdx.x & 63;
  if(row >= rows) return;
  double m = 0.0;
  for(long k = lane; k < Kin; k += 64) m = fmax(m, fabs(x[row * Kin + k]));
#pragma unroll
  for(int o = 32; o >= 1; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if(lane == 0) {
    int ex = kNoRow;
    if(m > 0.0) ex = 13 - ilogb(m);
    ex     = ex > kNoRow ? kNoRow : (ex < -kNoRow ? -kNoRow : ex);
    e[row] = (short) ex;
  }
}

// Row r = (x*D2 + y)*D3 + z of an operand array; each coordinate is an
// occupied (type 0) or virtual (type 1) orbital.  Its block is the tile
// triple of (x, y, z) in the same order.
struct BlockMap {
  long       D2, D3;
  int        n2, n3;                  // tiles of coordinates 2 and 3
  const int* t1;
  const int* t2;
  const int* t3;                      // orbital -> tile maps of the three coordinates
  __device__ long operator()(long r) const {
    const long z = r % D3, y = (r / D3) % D2, x = r / (D2 * D3);
    return ((long) t1[x] * n2 + t2[y]) * n3 + t3[z];
  }
};

__global__ void k_fill_i32(int* p, long n, int v) {
  const long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) p[i] = v;
}

// Block exponent = min over the block's rows (the largest row sets it).
__global__ void k_block_exp(const short* __restrict__ er, long rows, BlockMap bm, int* eb) {
  const long r = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if(r >= rows) return;
  const int e = er[r];
  if(e < kNoRow) atomicMin(&eb[bm(r)], e);
}

// Split: y = x * 2^e[block(row)] = hi + lo (FP16), written per 8 k as [hi8][lo8].
__global__ void k_split(const double* __restrict__ x, long rows, long Kin, long Kout, BlockMap bm,
                        const int* __restrict__ eb, _Float16* __restrict__ out) {
  const long n = rows * Kout;
  for(long i = (long) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long) gridDim.x * blockDim.x) {
    const long     row  = i / Kout, k = i % Kout;
    const double   y    = k < Kin ? ldexp(x[row * Kin + k], eb[bm(row)]) : 0.0;
    const _Float16 h    = (_Float16) y;
    const _Float16 l    = (_Float16) (y - (double) h);
    const long     base = row * Kout * 2 + (k / 8) * 16 + (k % 8);
    out[base]           = h;
    out[base + 8]       = l;
  }
}

// One thread per (task, part): the part's sign and term coefficients times
// 2^-(eA + eB), eA and eB the exponents of the (tile, tile, tile) blocks of
// the two operands (the same for every row of the part).
__global__ void k_part_scales(const TaskDev* __restrict__ tasks, int n_tasks, MixedDev M,
                              float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= 18 * n_tasks) return;
  const TaskDev& T  = tasks[i / 18];
  const int      gp = i % 18;
  if(!part_on(T.d1code, T.d2code, gp)) {
    out[i] = 0.f;
    return;
  }
  const unsigned code = part_code(T.d1code, T.d2code, gp);
  const int      g = gp >> 1, d2 = gp & 1, hs = g / 3, ps = g % 3;
  const int      hA = pair_a(hs), hB = pair_b(hs), pA = pair_a(ps), pB = pair_b(ps);
  const int      no = M.nto, nv = M.ntv;
  const int      tHS = T.tile[hs], tHA = T.tile[hA], tHB = T.tile[hB];
  const int      tPS = T.tile[3 + ps], tPA = T.tile[3 + pA], tPB = T.tile[3 + pB];
  int            aa, ka, ab, kb;
  float          ca, cb;
  if(!d2) {  // t2h blocks (V,V,O): (pA,pB,hs) / (pB,pA,hs); ovk (O,V,O): (hB,ps,hA) / (hA,ps,hB)
    choose_term(code, 4, 8, 1, 0, (tPA * nv + tPB) * no + tHS, (tPB * nv + tPA) * no + tHS, aa, ka, ca);
    choose_term(code, 16, 32, 3, 2, (tHB * nv + tPS) * no + tHA, (tHA * nv + tPS) * no + tHB, ab, kb, cb);
  }
  else {  // xv blocks (V,V,O): (pB,pA,hs) / (pA,pB,hs); t2p (O,O,V): (hA,hB,ps) / (hB,hA,ps)
    choose_term(code, 4, 8, 5, 4, (tPB * nv + tPA) * no + tHS, (tPA * nv + tPB) * no + tHS, aa, ka, ca);
    choose_term(code, 16, 32, 7, 6, (tHA * no + tHB) * nv + tPS, (tHB * no + tHA) * nv + tPS, ab, kb, cb);
  }
  const float sg = (code & 2) ? -1.f : 1.f;
  out[i]         = sg * ca * cb * ldexpf(1.f, -(M.eblk[aa][ka] + M.eblk[ab][kb]));
}

}  // namespace

int build_part_scales(const TaskDev* tasks, int n_tasks, const MixedDev& M, float* out,
                      void* stream) {
  if(n_tasks <= 0) return 0;
  hipLaunchKernelGGL(k_part_scales, dim3((unsigned) ((18l * n_tasks + 255) / 256)), dim3(256), 0,
                     (hipStream_t) stream, tasks, n_tasks, M, out);
  return (int) hipGetLastError();
}

int build_split(const SpatialBufs& B, int O, int V, const int* tile_o, int nto, const int* tile_v,
                int ntv, SplitBufs& S, void* stream) {
  hipStream_t   st = (hipStream_t) stream;
  const long    o = O, v = V, op = (O + 7) / 8 * 8, vp = (V + 7) / 8 * 8;
  const long    ko = (O + 15) / 16 * 16, kv = (V + 15) / 16 * 16;
  const double* src[kOperands] = {B.t2h, B.t2ha, B.ovk, B.ovka, B.xv, B.xva, B.t2p, B.t2pa};
  // row coordinates of each array (0 occupied, 1 virtual), slowest first
  const int types[kOperands][3] = {{1, 1, 0}, {1, 1, 0}, {0, 1, 0}, {0, 1, 0},
                                   {1, 1, 0}, {1, 1, 0}, {0, 0, 1}, {0, 0, 1}};
  int *      dto = nullptr, *dtv = nullptr;
  short*     er  = nullptr;
  hipError_t err = hipSuccess;
  auto       done = [&](hipError_t e) {
    if(dto) (void) hipFree(dto);
    if(dtv) (void) hipFree(dtv);
    if(er) (void) hipFree(er);
    return (int) e;
  };
  if((err = hipMalloc((void**) &dto, sizeof(int) * o)) || (err = hipMalloc((void**) &dtv, sizeof(int) * v)))
    return done(err);
  if((err = hipMemcpyAsync(dto, tile_o, sizeof(int) * o, hipMemcpyHostToDevice, st)) ||
     (err = hipMemcpyAsync(dtv, tile_v, sizeof(int) * v, hipMemcpyHostToDevice, st)))
    return done(err);
  {
    const size_t zb = 4 * (size_t) (ko > kv ? ko : kv);
    if((err = hipMalloc((void**) &S.zero, zb))) return done(err);
    if((err = hipMemsetAsync(S.zero, 0, zb, st))) return done(err);
  }
  const long rows_max = v * v * o > o * o * v ? v * v * o : o * o * v;
  // k_row_exp runs one wave per row: 64 * rows work-items in one dispatch
  if(rows_max * 64 > (long) UINT32_MAX) return done(hipErrorInvalidConfiguration);
  if((err = hipMalloc((void**) &er, sizeof(short) * rows_max))) return done(err);
  for(int j = 0; j < kOperands; j++) {
    const int* const tmap[2] = {dto, dtv};
    const long       dim[2]  = {o, v};
    const int        nt[2]   = {nto, ntv};
    const int*       ty      = types[j];
    BlockMap         bm;
    bm.D2 = dim[ty[1]];
    bm.D3 = dim[ty[2]];
    bm.n2 = nt[ty[1]];
    bm.n3 = nt[ty[2]];
    bm.t1 = tmap[ty[0]];
    bm.t2 = tmap[ty[1]];
    bm.t3 = tmap[ty[2]];
    const long rows   = dim[ty[0]] * dim[ty[1]] * dim[ty[2]];
    const long blocks = (long) nt[ty[0]] * nt[ty[1]] * nt[ty[2]];
    const long kin = j < 4 ? op : vp, kout = j < 4 ? ko : kv;
    S.rows[j]      = rows;
    if((err = hipMalloc((void**) &S.eblk[j], sizeof(int) * blocks))) return done(err);
    if((err = hipMalloc((void**) &S.v[j], 4 * rows * kout))) return done(err);
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned) ((blocks + 255) / 256)), dim3(256), 0, st,
                       S.eblk[j], blocks, kNoRow);
    hipLaunchKernelGGL(k_row_exp, dim3((unsigned) ((rows + 3) / 4)), dim3(256), 0, st, src[j], rows,
                       kin, er);
    hipLaunchKernelGGL(k_block_exp, dim3((unsigned) ((rows + 255) / 256)), dim3(256), 0, st, er,
                       rows, bm, S.eblk[j]);
    // grid-stride: a dispatch's grid size is a 32-bit count of work-items
    const long n = rows * kout;
    hipLaunchKernelGGL(k_split, dim3((unsigned) std::min<long>((n + 255) / 256, 1l << 20)), dim3(256), 0, st, src[j], rows,
                       kin, kout, bm, S.eblk[j], reinterpret_cast<_Float16*>(S.v[j]));
    if((err = hipGetLastError())) return done(err);
  }
  return done(hipStreamSynchronize(st));
}

int launch_spatial_mixed(const MixedDev& M, const RunDev& R, int grid, void* stream) {
  hipLaunchKernelGGL(k_spatial_mixed, dim3(grid), dim3(NT), 0, (hipStream_t) stream, M, R);
  return (int) hipGetLastError();
}

}  // namespace xt
