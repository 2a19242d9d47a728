// Device-side gather of the per-task packed operand buffers: the work of the
// reference's host gathers ccsd_t_data_s1_new (ccsd_t_all_fused_singles.hpp:34-391),
// ccsd_t_data_d1_new (ccsd_t_all_fused_doubles1.hpp:37-307) and ccsd_t_data_d2_new
// (ccsd_t_all_fused_doubles2.hpp:37-360), which read TAMM blocks (Tensor::get)
// and permute them with hptt, done here from the device-resident closed-shell
// inputs.  The host plans the list of blocks (xt_gather.cpp); this kernel
// writes every element of every block.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "xt_device.h"

namespace xt {
namespace {

// Spin-orbital element of one of the five tensors the gathers read, formed
// from the spatial closed-shell inputs (spin-orbital order is [alpha, beta]
// per space).  The combination rule c1*x - c2*y and the spatial sources:
//   T1   t1(a,i)        = [sa=si] t1[A][I]
//   T2   t2(a,b,i,j)    = [sa=si][sb=sj] t2[A][B][I][J] - [sa=sj][sb=si] t2[A][B][J][I]
//   IJAB <ij||ab>       = [si=sa][sj=sb] (ai|bj) - [si=sb][sj=sa] (bi|aj)          (vvoo)
//   IJKA <ij||ka>       = [si=sk][sj=sa] (ja|ik) - [sj=sk][si=sa] (ia|jk)          (ovoo)
//   IABC <ia||bc>       = [si=sb][sa=sc] (ac|bi) - [si=sc][sa=sb] (ab|ci)          (vvvo)
// with t2 read from t2h[a][b][i][Op], ovoo from ovk[i][a][j][Op] and vvvo
// from xv[c][b][i][Vp] = vvvo[a][c][b][i].
__device__ __forceinline__ double so_value(const GatherSrc& S, int kind, const int* sp, const long* x) {
  const long O = S.O, V = S.V, Op = S.Op, Vp = S.Vp;
  auto comb = [](bool c1, bool c2, double a, double b) {
    return c1 && c2 ? a - b : (c1 ? a : (c2 ? -b : 0.0));
  };
  switch(kind) {
    case kGatherT1: return sp[0] == sp[1] ? S.t1[x[0] * O + x[1]] : 0.0;
    case kGatherT2: {
      const bool c1 = sp[0] == sp[2] && sp[1] == sp[3], c2 = sp[0] == sp[3] && sp[1] == sp[2];
      const long ab = (x[0] * V + x[1]) * O;
      return comb(c1, c2, c1 ? S.t2h[(ab + x[2]) * Op + x[3]] : 0.0,
                  c2 ? S.t2h[(ab + x[3]) * Op + x[2]] : 0.0);
    }
    case kGatherIJAB: {
      const bool c1 = sp[0] == sp[2] && sp[1] == sp[3], c2 = sp[0] == sp[3] && sp[1] == sp[2];
      return comb(c1, c2, c1 ? S.vvoo[((x[2] * V + x[3]) * O + x[0]) * O + x[1]] : 0.0,
                  c2 ? S.vvoo[((x[3] * V + x[2]) * O + x[0]) * O + x[1]] : 0.0);
    }
    case kGatherIJKA: {
      const bool c1 = sp[0] == sp[2] && sp[1] == sp[3], c2 = sp[1] == sp[2] && sp[0] == sp[3];
      return comb(c1, c2, c1 ? S.ovk[((x[1] * V + x[3]) * O + x[0]) * Op + x[2]] : 0.0,
                  c2 ? S.ovk[((x[0] * V + x[3]) * O + x[1]) * Op + x[2]] : 0.0);
    }
    default: {  // kGatherIABC
      const bool c1 = sp[0] == sp[2] && sp[1] == sp[3], c2 = sp[0] == sp[3] && sp[1] == sp[2];
      return comb(c1, c2, c1 ? S.xv[((x[3] * V + x[2]) * O + x[0]) * Vp + x[1]] : 0.0,
                  c2 ? S.xv[((x[2] * V + x[3]) * O + x[0]) * Vp + x[1]] : 0.0);
    }
  }
}

// Source dimension whose consecutive values are adjacent in memory, for the
// term the block reads (for a T2 block that reads both terms, the first one).
__device__ __forceinline__ int fast_dim(const GatherBlock& B) {
  switch(B.kind) {
    case kGatherIJKA: return 2;
    case kGatherT2: return (B.spin[0] == B.spin[2] && B.spin[1] == B.spin[3]) ? 3 : 2;
    default: return 1;  // T1, IJAB, IABC
  }
}

constexpr int GT   = 256;      // threads per gather workgroup
constexpr int SLAB = 48 * 49;  // LDS slab (doubles): extents up to 48 with one pad column

// blockIdx.y = block.  Destination element e has (row-major, last fastest)
// coordinates i[k] over the permuted extents size[perm[k]]; source coordinate
// perm[k] is i[k] (hptt's row-major permutation), and the value is alpha
// times the element.
//
// Where the destination's fastest dimension is also the source's contiguous
// one, threads stride over destination elements directly.  Otherwise the block
// goes through LDS in slabs over (destination position of the source's
// contiguous dimension) x (destination's fastest position), one slab per pair
// of the two other positions: the slab is read along the source's contiguous
// dimension and written along the destination's, so both sides run through
// whole rows instead of one element per row.
__global__ __launch_bounds__(GT) void k_gather(const GatherBlock* __restrict__ blocks, GatherSrc S,
                                               GatherDst D) {
  __shared__ double  slab[SLAB];
  const GatherBlock& B   = blocks[blockIdx.y];
  double*            dst = D.buf[B.buf] + B.dst;
  const int          nd  = B.nd;
  const unsigned     n   = (unsigned) B.n;  // a block holds fewer than 2^31 doubles (host check)
  int                sp[4] = {0, 0, 0, 0};
#pragma unroll
  for(int d = 0; d < 4; d++) sp[d] = d < nd ? B.spin[d] : 0;
  // destination positions padded to four (leading extents 1): source
  // dimension src[k] (-1 for padding) and extent E[k]
  int src[4], E[4];
#pragma unroll
  for(int k = 0; k < 4; k++) {
    const int kk = k - (4 - nd);
    src[k]       = kk >= 0 ? B.perm[kk] : -1;
    E[k]         = kk >= 0 ? B.size[src[k]] : 1;
  }
  const int fd = fast_dim(B);
  int       kS = 3;
#pragma unroll
  for(int k = 0; k < 3; k++)
    if(src[k] == fd) kS = k;
  const unsigned ES = (unsigned) E[kS], EF = (unsigned) E[3], ld = EF + 1;
  if(kS == 3 || ES * ld > (unsigned) SLAB) {
    // 32-bit index decode: a 64-bit divide costs several times more
    for(unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
      unsigned r = e;
      long     x[4] = {0, 0, 0, 0};
      for(int k = nd - 1; k >= 0; k--) {
        const int      d  = B.perm[k];
        const unsigned sz = (unsigned) B.size[d];
        x[d]              = B.start[d] + (long) (r % sz);
        r /= sz;
      }
      dst[e] = B.alpha * so_value(S, B.kind, sp, x);
    }
    return;
  }
  // the two other positions o1 < o2 index the slabs
  const int o1 = kS == 0 ? 1 : 0, o2 = kS == 2 ? 1 : 2;
  unsigned  st[4];
  st[3] = 1;
#pragma unroll
  for(int k = 2; k >= 0; k--) st[k] = st[k + 1] * (unsigned) E[k + 1];
  const unsigned E2 = (unsigned) E[o2], nslab = (unsigned) E[o1] * E2, nel = ES * EF;
  const int      d1 = src[o1], d2 = src[o2], dF = src[3];
  long           x0[4];
#pragma unroll
  for(int d = 0; d < 4; d++) x0[d] = d < nd ? B.start[d] : 0;
  // every thread of the workgroup runs the same slab loop, so the barriers
  // below are reached by all of them
  for(unsigned sl = blockIdx.x; sl < nslab; sl += gridDim.x) {
    const unsigned a = sl / E2, b = sl - a * E2;
    __syncthreads();  // the previous slab's LDS reads are done
    for(unsigned e = threadIdx.x; e < nel; e += GT) {
      const unsigned f = e / ES, s = e - f * ES;
      long           x[4];
#pragma unroll
      for(int d = 0; d < 4; d++)
        x[d] = x0[d] + (d == fd ? s : d == dF ? f : d == d1 ? a : d == d2 ? b : 0u);
      slab[s * ld + f] = B.alpha * so_value(S, B.kind, sp, x);
    }
    __syncthreads();
    const unsigned base = a * st[o1] + b * st[o2];
    for(unsigned e = threadIdx.x; e < nel; e += GT) {
      const unsigned s = e / EF, f = e - s * EF;
      dst[base + s * st[kS] + f] = slab[s * ld + f];
    }
  }
}

// Dense spin-orbital amplitudes, [alpha, beta] per space:
//   t1_so[A][I] (2V x 2O), t2_so[A][B][I][J] ((2V)^2 (2O)^2), A = spin * V + a.
// blockIdx.y strides over rows (A for T1, the (A, B) pair for T2), threads
// over the occupied index (pair) of the row.
__global__ void k_expand_t(GatherSrc S, double* __restrict__ t1_so, double* __restrict__ t2_so) {
  const long nv = 2 * S.V, no = 2 * S.O;
  if(t1_so)
    for(long a = blockIdx.y; a < nv; a += gridDim.y)
      for(long i = (long) blockIdx.x * blockDim.x + threadIdx.x; i < no; i += (long) gridDim.x * blockDim.x) {
        const long x[4]  = {a % S.V, i % S.O, 0, 0};
        const int  sp[4] = {(int) (a / S.V), (int) (i / S.O), 0, 0};
        t1_so[a * no + i] = so_value(S, kGatherT1, sp, x);
      }
  if(t2_so) {
    // spin = index >= n, spatial = index - spin * n: no divide per element
    // but the 32-bit one that splits the occupied pair
    const unsigned uo = (unsigned) S.O, uno = (unsigned) no, nq = uno * uno;
    for(long p = blockIdx.y; p < nv * nv; p += gridDim.y) {
      const long a = p / nv, b = p % nv;
      const int  sa = a >= S.V, sb = b >= S.V;
      double*    row = t2_so + p * (long) nq;
      for(unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += gridDim.x * blockDim.x) {
        const unsigned i = q / uno, j = q - i * uno;
        const int      si = i >= uo, sj = j >= uo;
        const long     x[4]  = {a - sa * S.V, b - sb * S.V, (long) (i - si * uo), (long) (j - sj * uo)};
        const int      sp[4] = {sa, sb, si, sj};
        row[q]               = so_value(S, kGatherT2, sp, x);
      }
    }
  }
}

}  // namespace

int launch_expand_t(const GatherSrc& S, double* t1_so, double* t2_so, void* stream) {
  const long     no = 2 * S.O, nv = 2 * S.V;
  const unsigned gx = (unsigned) std::min<long>((no * no + 255) / 256, 64);
  const unsigned gy = (unsigned) std::min<long>(nv * nv, 65535);
  hipLaunchKernelGGL(k_expand_t, dim3(gx, gy), dim3(256), 0, (hipStream_t) stream, S, t1_so, t2_so);
  return (int) hipGetLastError();
}

int launch_gather(const GatherBlock* blocks, int n_blocks, long max_elems, const GatherSrc& S,
                  const GatherDst& D, void* stream) {
  if(n_blocks <= 0) return 0;
  if(n_blocks > 65535) return (int) hipErrorInvalidConfiguration;
  const long     per = (max_elems + 255) / 256;
  const unsigned gx  = (unsigned) (per < 512 ? (per > 0 ? per : 1) : 512);
  hipLaunchKernelGGL(k_gather, dim3(gx, (unsigned) n_blocks), dim3(GT), 0, (hipStream_t) stream,
                     blocks, S, D);
  return (int) hipGetLastError();
}

}  // namespace xt
