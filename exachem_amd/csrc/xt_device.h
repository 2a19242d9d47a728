// Device-side data structures and launch entry points shared by the host
// planner (xt_host.cpp) and the gfx950 kernels (xt_kernels.hip).
#pragma once
#include <cstddef>
#include <cstdint>

namespace xt {

// One energy task (h1b,h2b,h3b,p4b,p5b,p6b) of the fused driver
// (ccsd_t_fused_driver.hpp:368-395) as the kernels see it.  Positions 0..5 are
// h1,h2,h3,p4,p5,p6.
struct TaskDev {
  int       dim[6];      // tile extents
  int       off[6];      // first spatial orbital of each tile (occ or virt range)
  int       tile[6];     // spatial tile index (occupied or virtual tile list) per position
  int       nsub[6];     // ceil(dim/4): 4^6 sub-tiles of the FP64 kernel per position
  int       nsubm[6];    // mixed kernel (4,4,4 | 8,8,8) boxes: ceil(dim/4), ceil(dim/8)
  int       n_sub;       // product of nsub
  // Per output bipartition g = 3*hs + ps (hs: single hole position 0..2,
  // ps: single particle position 0..2), 6 bits each:
  //   bit0 enabled, bit1 negative sign, bit2 a1, bit3 a2, bit4 b1, bit5 b2
  // (a*: M-side operand terms, b*: N-side operand terms).
  uint64_t d1code;
  uint64_t d2code;
  // s1 slot s = 3*pa + hi, 4 bits each: bit0 enabled, bit1 negative, bit2 b1, bit3 b2.
  uint64_t s1code;
  double   factor;
  double   flops;       // analytic flop count of the task
};

// HBM-resident closed-shell inputs in kernel layouts.  The contraction index
// (last) of every operand array is padded with zeros to Op = roundup(O, 8)
// or Vp = roundup(V, 8), so K-chunks load with aligned vector loads.  The *a
// arrays hold the antisymmetrised difference used by same-spin blocks, formed
// once in FP64 exactly as the spin-orbital element would be (x - y).
struct SpatialDev {
  int           O, V, Op, Vp;
  const double* eo;    // [O]
  const double* ev;    // [V]
  const double* t1;    // [V][O]           t_i^a
  const double* t2h;   // [V][V][O][Op]    t2h[a][b][i][k] = t_ik^ab            (d1 M side)
  const double* t2ha;  // [V][V][O][Op]    t2h[a][b][i][k] - t2h[b][a][i][k]
  const double* ovk;   // [O][V][O][Op]    ovk[i][a][j][k] = (ia|jk)            (d1 N side)
  const double* ovka;  // [O][V][O][Op]    ovk[i][a][j][k] - ovk[j][a][i][k]
  const double* xv;    // [V][V][O][Vp]    xv[c][b][i][a] = (ac|bi)             (d2 M side)
  const double* xva;   // [V][V][O][Vp]    xv[c][b][i][a] - xv[b][c][i][a]
  const double* t2p;   // [O][O][V][Vp]    t2p[i][j][a][b] = t_ij^ab            (d2 N side)
  const double* t2pa;  // [O][O][V][Vp]    t2p[i][j][a][b] - t2p[j][i][a][b]
  const double* vvoo;  // [V][V][O][O]     (ai|bj)                              (s1)
};

// Device copies of the inputs in the layouts above.
struct SpatialBufs {
  double *eo, *ev, *t1, *t2h, *t2ha, *ovk, *ovka, *xv, *xva, *t2p, *t2pa, *vvoo;
};

// Split FP16 operands of the mixed-precision kernel.  Array j (same order as
// the FP64 operand arrays: 0 t2h, 1 t2ha, 2 ovk, 3 ovka, 4 xv, 5 xva, 6 t2p,
// 7 t2pa) holds, for each contraction run (row) r, the values
// y = x * 2^e[block(r)] split as y ~ hi + lo (FP16 each), stored per 8
// consecutive k as [hi x8][lo x8]; rows are padded to Kp16 = roundup(K, 16)
// with zeros.  A block is the set of rows whose three row coordinates fall in
// one (tile, tile, tile) triple of the tiling; its exponent puts the block's
// largest |y| in [2^13, 2^14) (per-tile amplitude scaling), so the scale of a
// whole contraction part of a task is one number.
constexpr int kOperands = 8;
struct SplitBufs {
  unsigned char* v[kOperands];     // [rows][Kp16/8][2][8] halves
  int*           eblk[kOperands];  // [blocks] exponents, row-coordinate tile order
  long           rows[kOperands];
  unsigned char* zero;             // zero row, 4 * max(Kp16) bytes
  float*         pscale;           // [task][18] contraction part scales (0 for parts off)
  long           pscale_n;         // tasks pscale was built for
};
struct MixedDev {
  int                  O, V, Ko, Kv;  // Ko = roundup(O,16), Kv = roundup(V,16)
  int                  nto, ntv;      // number of occupied / virtual spatial tiles
  const double*        eo;
  const double*        ev;
  const double*        t1;
  const double*        vvoo;
  const unsigned char* v[kOperands];
  const int*           eblk[kOperands];
  const unsigned char* zero;  // zero row (4 * max(Ko, Kv) bytes): rows outside a tile read it
  const float*         pscale;  // [task][18] part scales (build_part_scales)
};

// Work queues: one per XCD.  Queue q holds the tasks qtask[qbeg[q]..qbeg[q+1])
// in processing order; entry e owns a contiguous ticket range of its queue
// starting at qtick[e].  A ticket is one box of the kernel (one 4^6 sub-tile
// for FP64, one (4,4,4|8,8,8) box for mixed).  64 consecutive tickets form a
// group of neighbouring boxes, 2^glog[p] along position p (sum of glog = 6):
// the group shape decides which operand panels the workgroups resident on one
// XCD share in its L2.  A task has 64 * prod(ceil(nsub[p] / 2^glog[p]))
// tickets; tickets outside the task's box grid are skipped.
constexpr int kQueues = 8;
// Threads of the one workgroup that sums the task energies in a fixed order
// (k_reduce_final); xt_sum_task_energies repeats the same order on the host.
constexpr int kReduceThreads = 256;
constexpr int kGroupLog2 = 6;  // 64 boxes per ticket group

struct RunDev {
  const TaskDev*      tasks;
  const int*          qtask;   // [n_entries]
  const long long*    qtick;   // [n_entries]
  const int*          qbeg;    // [kQueues+1]
  const long long*    qtotal;  // [kQueues]
  unsigned long long* qctr;    // [kQueues] ticket counters (zeroed before a run)
  unsigned long long* err;     // [1] set by a box whose energy partial is not finite or
                               // beyond 2^62 Eh (zeroed before a run, read after it)
  unsigned long long* acc;     // [4*n_tasks] (E4 lo, E4 hi, E5 lo, E5 hi): 128-bit
                               // two's-complement fixed point, scale 2^64 per Eh
  unsigned long long* cost;    // [n_tasks] summed workgroup wall-clock ticks (optional)
  unsigned long long* prog;    // [kQueues] host-mapped mirror of the ticket counters (optional)
  int                 n_tasks;
  int                 glog[6];  // ticket group shape (log2 extent per position, sum 6)
};

// Device allocations of the upload paths (xt_host.cpp): hipMalloc / hipFree,
// also counted (live bytes and their maximum) while the calling thread's
// upload has a count open, so that the upload's peak can be compared with the
// memory guard's estimate (xt_upload_peak_bytes).  Returns hipError_t as int.
int  track_malloc(void** p, size_t bytes);
void track_free(void* p);

// Kernel launchers (defined in xt_kernels.hip).  Return hipError_t as int.
// Builds the padded / antisymmetrised operand arrays from the caller's
// device copies t2 [V][V][O][O], ovoo [O][V][O][O] and vvvo [V][V][V][O]
// (consumed; vvvo is freed as soon as xv exists to bound peak memory).
// Returns 0 or a hipError_t / -1 for allocation failure.
// vvoo [V][V][O][O], ovoo [O][V][O][O] and xv (the vvvo operand layout) from
// Cholesky vectors L [O+V][O+V][K] on the device (xt_chol.cpp, rocBLAS).
int cholesky_integrals(int O, int V, long K, const double* L, double* vvoo, double* ovoo,
                       double* xv, void* stream);
// xva = xv with the (c,b) pair antisymmetrised (xv must exist).
int build_xva(int O, int V, SpatialBufs& B, void* stream);
// Reorders the (ai|bj) GEMM result [(a,i)][(b,j)] into vvoo[a][b][i][j].
int launch_vvoo_order(const double* in, double* out, int O, int V, void* stream);
// With vvvo == nullptr the xv/xva pair is left to the caller.
int build_operands(int O, int V, double* t2, double* ovoo, double*& vvvo, SpatialBufs& B,
                   void* stream);
int launch_spatial_fp64(const SpatialDev& S, const RunDev& R, int grid, void* stream);
int launch_reduce(const TaskDev* tasks, int n_tasks, const unsigned long long* acc,
                  double* task_e, double* e45, void* stream);
int launch_zero_u64(unsigned long long* p, long n, void* stream);
// tile_o[O] / tile_v[V]: spatial tile index of every orbital (host arrays).
int build_split(const SpatialBufs& B, int O, int V, const int* tile_o, int nto, const int* tile_v,
                int ntv, SplitBufs& S, void* stream);
// Device gather of the per-task packed buffers (xt_gather.hip, xt_gather.cpp).
// One GatherBlock is one Tensor::get + hptt permute of the host gathers: the
// source block's extents, spins and spatial starts per dimension (source
// order), the permutation (destination dimension k is source dimension
// perm[k], row-major), the factor, and where it lands.
enum GatherKind { kGatherT1 = 0, kGatherT2 = 1, kGatherIJAB = 2, kGatherIJKA = 3, kGatherIABC = 4 };
struct GatherBlock {
  long   dst;       // element offset in the destination buffer
  long   n;         // elements
  int    buf;       // destination: 0 d1_t2, 1 d1_v2, 2 d2_t2, 3 d2_v2, 4 s1_t1, 5 s1_v2
  int    kind;      // GatherKind
  int    nd;        // 2 (T1) or 4
  int    size[4];   // source extents
  int    start[4];  // spatial start per source dimension
  int    spin[4];   // 0 alpha, 1 beta
  int    perm[4];
  double alpha;     // +1 or -1
};
struct GatherSrc {
  long          O, V, Op, Vp;
  const double* t1;    // [V][O]
  const double* t2h;   // [V][V][O][Op]
  const double* vvoo;  // [V][V][O][O]
  const double* ovk;   // [O][V][O][Op]
  const double* xv;    // [V][V][O][Vp], xv[c][b][i][a] = vvvo[a][c][b][i]
};
struct GatherDst {
  double* buf[6];
};
int launch_gather(const GatherBlock* blocks, int n_blocks, long max_elems, const GatherSrc& S,
                  const GatherDst& D, void* stream);
// Dense spin-orbital T1 / T2 from the spatial amplitudes (either may be null).
int launch_expand_t(const GatherSrc& S, double* t1_so, double* t2_so, void* stream);

int launch_spatial_mixed(const MixedDev& M, const RunDev& R, int grid, void* stream);
// Scale of every contraction part of every task (sign and term coefficients
// times 2^-(eA + eB) of the blocks the part reads), from the split's block
// exponents: out[18 * task + gp].
int build_part_scales(const TaskDev* tasks, int n_tasks, const MixedDev& M, float* out,
                      void* stream);

// Task cost model (xt_mlp.hip): 7 inputs (6 tile extents / 48, precision
// flag) -> 64 -> 64 -> 1, FP32.
constexpr int kMlpIn     = 7;
constexpr int kMlpHidden = 64;
constexpr int kMlpParams = kMlpIn * kMlpHidden + kMlpHidden + kMlpHidden * kMlpHidden +
                           kMlpHidden + kMlpHidden + 1;
// Device buffers of one training batch of n samples.
struct MlpWork {
  float *P, *G, *M1, *M2;   // parameters, gradient, Adam moments   [kMlpParams]
  float *X, *T, *Y;         // inputs [n][7], targets [n], outputs [n]
  float *A1, *A2, *D1, *D2; // activations and error signals        [n][64]
  float *DY, *SQ;           // [n]
  float* loss;              // [1] mean squared error of the last pass
  int*   step;              // [2] Adam step count, its value at the start of training
};
int mlp_forward(const float* P, const float* X, long n, float* Y, float* A1, float* A2,
                void* stream);
int mlp_loss_grad(const MlpWork& W, long n, void* stream);
int mlp_train_step(const MlpWork& W, long n, float lr, float* loss_hist, void* stream);

// Tile-count limits of the per-task launcher (the reference's, nontc:56-57);
// the host rejects larger noab / nvab before any launch.
constexpr int kMaxNoab = 50, kMaxNvab = 140;
// The launcher's packed metadata, in ints: the host arrays d1_size[7*noab],
// d1_exec[9*noab], d2_size[7*nvab], d2_exec[9*nvab], s1_size[6], s1_exec[9],
// then the K tile lists the host builds from them: kKlistCountInts counts
// (one per bipartition: entries | d1 entries << 16, then padding to a
// 16-byte boundary) and, for bipartition g = 0..8 in turn, entries + 1 int4
// (family 0 d1 / 1 d2, exec slot, K, first packed k; the last entry holds the
// total K in .w).  The first d2 tile starts on a multiple of kPackedKC, the
// kernel's K chunk, so that every chunk reads one family.
// kPackedMetaInts is the size at the reference's tile-count limits.
constexpr int kPackedKC       = 16;
constexpr int kKlistCountInts = 9;  // 15 + 9 ints after the size/exec arrays: 16-byte aligned
constexpr int kPackedMetaInts =
  16 * kMaxNoab + 16 * kMaxNvab + 15 + kKlistCountInts + 4 * 9 * (kMaxNoab + kMaxNvab + 1);
static_assert((15 + kKlistCountInts) % 4 == 0, "tile lists must start on a 16-byte boundary");
static_assert(kPackedMetaInts % 4 == 0, "metadata slots must keep 16-byte alignment");
struct PackedArgs {
  int           dim[6];
  const double* d1_t2;
  const double* d1_v2;
  const double* d2_t2;
  const double* d2_v2;
  const double* s1_t1;
  const double* s1_v2;
  const int*    meta;  // device copy: d1_size[7*noab] d1_exec[9*noab] d2_size[7*nvab] d2_exec[9*nvab] s1_size[6] s1_exec[9]
  int           noab, nvab;
  long          max_d1_t2, max_d1_v2, max_d2_t2, max_d2_v2, max_s1_t1, max_s1_v2;
  const double* evl[6];
  double*       partial;  // [2*num_blocks]
  long          num_blocks;
};
int launch_packed_fp64(const PackedArgs& A, void* stream);
// The same launch without allocating: the n_meta ints of host_meta go into a
// fixed device array of kPackedMetaInts ints (one per GPU, kept until the
// process ends) through the arguments of small copy kernels on the stream,
// and A.meta is set to that array; after_copy (a hipEvent_t, may be null) is
// recorded between the copies and the launch.  Calls that overlap on one GPU
// race on the array.
int launch_packed_fp64_noalloc(PackedArgs& A, const int* host_meta, int n_meta, void* stream,
                               void* after_copy);

}  // namespace xt
