/*
 * exachem_t.h -- C ABI of the MI355X-native closed-shell CCSD(T) triples
 * energy path (library: exachem_amd/_lib/libexachem_t.so).
 *
 * Two call surfaces:
 *
 *  1. Per-task launcher, a drop-in for ExaChem's
 *       fully_fused_ccsd_t_gpu<double>(...)
 *     (exachem/cc/ccsd_t/ccsd_t_all_fused_nontcCuda_Hip_Sycl.cpp:2800-2817,
 *      declared at exachem/cc/ccsd_t/ccsd_t_all_fused.hpp:35-53).  It takes the
 *     reference's packed device buffers (built by ccsd_t_data_{s1,d1,d2}_new)
 *     and host size/exec arrays and writes 2*num_blocks partial energies that
 *     the caller sums and scales by the task factor, exactly like
 *     hostEnergyReduce (ccsd_t_all_fused.hpp:19-32).
 *
 *  2. Whole-problem path replacing the fused task loop
 *       CCSD_T_Fused_Driver<T>::execute(...)
 *     (exachem/cc/ccsd_t/ccsd_t_fused_driver.hpp:91-546): closed-shell spatial
 *     inputs resident in HBM, the reference's task list and factors, all tasks
 *     in one persistent launch, E[4]/E[5] out with a fixed-order FP64 sum.
 *
 * Every function returns XT_OK (0) or a negative XT_E* code and never exits;
 * xt_last_error() gives the message of the last failure on the calling thread.
 * No HIP or torch types appear in the signatures: streams and events are
 * passed as void* (hipStream_t / hipEvent_t or NULL).
 */
#ifndef EXACHEM_T_H
#define EXACHEM_T_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XT_OK 0
#define XT_EINVAL (-1)    /* bad argument (null pointer, bad size, limit exceeded)   */
#define XT_ENOMEM (-2)    /* device or host allocation failed                         */
#define XT_EHIP (-3)      /* HIP runtime error                                        */
#define XT_ESTATE (-4)    /* call out of order (e.g. run before upload)               */
#define XT_ENODEV (-5)    /* no usable gfx950 device                                  */
#define XT_ERANGE (-6)    /* a non-finite input, or an energy partial that is not finite
                             or beyond 2^62 Eh (e.g. a zero orbital-energy denominator):
                             the run's energies are not returned (set to NaN)          */

/* Precision of the contractions. */
#define XT_PREC_FP64 0   /* FP64 MFMA (v_mfma_f64_16x16x4f64), FP64 everywhere            */
#define XT_PREC_MIXED 1  /* FP16 hi/lo split products (hi*hi + hi*lo + lo*hi) on FP16 MFMA
                            with per-tile power-of-two scaling, FP32 accumulation, FP64
                            denominators, energy partials and sums                         */

/* Task ordering of the device work queue. */
#define XT_SCHED_DRIVER 0   /* reference driver order (ccsd_t_fused_driver.hpp:368-383)   */
#define XT_SCHED_ANALYTIC 1 /* longest-first by the analytic flop model                    */
#define XT_SCHED_LEARNED 2  /* longest-first by the learned cost model (falls back to 1)   */

/* ---------------------------------------------------------------- errors */
/* Message of the last failing call on this thread ("" if none). */
const char* xt_last_error(void);
/* Library version string. */
const char* xt_version(void);
/* 1 if a gfx950 device is visible, 0 otherwise (never initialises more than HIP). */
int xt_device_available(void);

/* ------------------------------------------------- 1. per-task launcher */
/*
 * Replaces fully_fused_ccsd_t_gpu<double> (nontc:2800-2817).  Argument order
 * and meaning follow the reference; stream/done_copy are hipStream_t /
 * hipEvent_t (done_copy may be NULL).  The host arrays host_d1_size[7*noab],
 * host_d1_exec[9*noab], host_d2_size[7*nvab], host_d2_exec[9*nvab],
 * host_s1_size[6], host_s1_exec[9] are consumed before the call returns, so
 * the caller may overwrite them at once (done_copy is recorded on return for
 * callers that wait on it).  partial_energies is device memory of
 * 2*num_blocks doubles, num_blocks = prod(ceil(base_size_x/4)).
 * Limits of the reference (MAX_NOAB=50, MAX_NVAB=140, nontc:56-57) are
 * checked and reported as XT_EINVAL instead of overflowing, as are tile
 * extents above 4096 and exec slot indices of 2^30 or more.
 */
int xt_fully_fused_ccsd_t_gpu(void* stream, size_t num_blocks, size_t base_size_h1b,
                              size_t base_size_h2b, size_t base_size_h3b, size_t base_size_p4b,
                              size_t base_size_p5b, size_t base_size_p6b,
                              const double* df_dev_d1_t2_all, const double* df_dev_d1_v2_all,
                              const double* df_dev_d2_t2_all, const double* df_dev_d2_v2_all,
                              const double* df_dev_s1_t1_all, const double* df_dev_s1_v2_all,
                              const int* host_d1_size, const int* host_d1_exec,
                              const int* host_d2_size, const int* host_d2_exec,
                              const int* host_s1_size, const int* host_s1_exec, size_t size_noab,
                              size_t size_max_dim_d1_t2, size_t size_max_dim_d1_v2,
                              size_t size_nvab, size_t size_max_dim_d2_t2,
                              size_t size_max_dim_d2_v2, size_t size_max_dim_s1_t1,
                              size_t size_max_dim_s1_v2, const double* dev_evl_sorted_h1b,
                              const double* dev_evl_sorted_h2b, const double* dev_evl_sorted_h3b,
                              const double* dev_evl_sorted_p4b, const double* dev_evl_sorted_p5b,
                              const double* dev_evl_sorted_p6b, double* partial_energies,
                              void* done_copy);

/* The per-task launcher above without any allocation: no counterpart in the
 * reference.  It is for callers that cannot allocate device memory, such as
 * code under HIP graph capture, tight serving loops or harnesses that forbid
 * hipMalloc.  Same arguments, meaning, checks and results (bit for bit) as
 * xt_fully_fused_ccsd_t_gpu; it calls no hipMalloc, hipFree or other
 * allocating routine.  The launcher's metadata (the host size/exec arrays
 * and the K tile lists built from them) goes into one fixed device array per
 * GPU of 9,940 ints (39,760 bytes: its size at the reference's MAX_NOAB = 50
 * / MAX_NVAB = 140 limits, larger noab / nvab fail with XT_EINVAL), held
 * until the process ends and written by small copy kernels
 * on `stream` whose arguments carry the host arrays, so the caller may reuse
 * them when the call returns (done_copy is recorded after those copies).
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (different streams or host threads, without ordering between them) race. */
int xt_fully_fused_ccsd_t_gpu_noalloc(
  void* stream, size_t num_blocks, size_t base_size_h1b, size_t base_size_h2b,
  size_t base_size_h3b, size_t base_size_p4b, size_t base_size_p5b, size_t base_size_p6b,
  const double* df_dev_d1_t2_all, const double* df_dev_d1_v2_all,
  const double* df_dev_d2_t2_all, const double* df_dev_d2_v2_all,
  const double* df_dev_s1_t1_all, const double* df_dev_s1_v2_all, const int* host_d1_size,
  const int* host_d1_exec, const int* host_d2_size, const int* host_d2_exec,
  const int* host_s1_size, const int* host_s1_exec, size_t size_noab, size_t size_max_dim_d1_t2,
  size_t size_max_dim_d1_v2, size_t size_nvab, size_t size_max_dim_d2_t2,
  size_t size_max_dim_d2_v2, size_t size_max_dim_s1_t1, size_t size_max_dim_s1_v2,
  const double* dev_evl_sorted_h1b, const double* dev_evl_sorted_h2b,
  const double* dev_evl_sorted_h3b, const double* dev_evl_sorted_p4b,
  const double* dev_evl_sorted_p5b, const double* dev_evl_sorted_p6b, double* partial_energies,
  void* done_copy);

/* --------------------------------------------- 2. whole-problem path */
typedef struct xt_ctx xt_ctx;

/* One context per host thread; no global mutable state is shared.  device = -1
 * creates a host-only context (tiling, task list, flop count; no HIP calls). */
int xt_create(int device, xt_ctx** out);
int xt_destroy(xt_ctx* ctx);

/*
 * Tiling of the spin-orbital space |occ_a|occ_b|virt_a|virt_b| as built by
 * setupMOIS (exachem/cholesky/cholesky_2e.cpp:186-277): the beta tiles mirror
 * the alpha tiles; k_spin is 1 for the first half of each space and 2 for the
 * second (exachem/cc/ccsd_t/ccsd_t.cpp:245-249).  occ_tiles[n_occ_tiles] sum
 * to O and vir_tiles[n_vir_tiles] sum to V.  Builds the task list of the fused
 * driver (ccsd_t_fused_driver.hpp:368-395) with its spin filter and factors.
 */
int xt_set_tiling(xt_ctx* ctx, int O, int V, int n_occ_tiles, const int* occ_tiles,
                  int n_vir_tiles, const int* vir_tiles, int is_restricted);

/* Number of energy tasks after the spin filter. */
int xt_num_tasks(xt_ctx* ctx, long* n_out);
/* tasks[6*n] = (h1b,h2b,h3b,p4b,p5b,p6b) in driver order, factors[n]. */
int xt_task_list(xt_ctx* ctx, long* tasks, double* factors);
/* Analytic flop count of helper_calculate_num_ops (fused_common.hpp:131-261):
 * per_task[n] (optional) and the total (exact integer held in a double up to 2^53,
 * also returned split as s1/d1/d2 in parts[3], optional). */
int xt_flops(xt_ctx* ctx, double* per_task, double* total, double* parts);

/*
 * Closed-shell spatial inputs (host memory, row-major FP64):
 *   eps_occ[O], eps_vir[V]      orbital energies (diagonal of F)
 *   t1[V][O]                    t_i^a
 *   t2[V][V][O][O]              t_ij^ab  (t_ij^ab = t_ji^ba)
 *   vvoo[V][V][O][O]            (ai|bj)  ((ai|bj) = (bj|ai))
 *   ovoo[O][V][O][O]            (ia|jk)  (symmetric in j,k)
 *   vvvo[V][V][V][O]            (ab|ci)  (symmetric in a,b)
 * Copied to HBM and re-laid out there for the kernels.  Must follow
 * xt_set_tiling with the same O, V.  Every element must be finite: the
 * upload checks that before any device work and fails with XT_ERANGE.
 */
/* Same inputs with the V2 blocks built on the device from Cholesky vectors,
 * in place of V2TensorSetup::setupV2Tensors (cholesky/v2tensors.cpp:52-98):
 * (pq|rs) = sum_Q chol[p][q][Q] * chol[r][s][Q] over spatial orbitals p, q,
 * r, s in [0, O+V) (occupied first); chol is [O+V][O+V][nchol], the order of
 * cholVpr(p, q, cind), symmetric in p, q.  eps_occ, eps_vir, t1, t2 as in
 * xt_upload_closed_shell.  The GEMMs over the Cholesky index run on the GPU
 * (rocBLAS); the V^3 O block never exists on the host, and the largest
 * device temporary besides the vectors is V^2 O^2.  (O+V) * nchol and V*O
 * must stay below 2^31 (BLAS leading dimensions). */
int xt_upload_cholesky(xt_ctx* ctx, const double* eps_occ, const double* eps_vir,
                       const double* t1, const double* t2, long nchol, const double* chol);
int xt_upload_closed_shell(xt_ctx* ctx, const double* eps_occ, const double* eps_vir,
                           const double* t1, const double* t2, const double* vvoo,
                           const double* ovoo, const double* vvvo);

/* Seeded synthetic closed-shell inputs, generated on the host by a
 * counter-based generator (independent of thread count) and uploaded as by
 * xt_upload_closed_shell.  Distributions: eps_occ ~ U[-2,-0.3] sorted,
 * eps_vir ~ U[0.1,3] sorted, vvoo/ovoo/vvvo ~ N(0, sigma^2) with exact
 * permutational symmetry, t2 = vvoo/(e_i+e_j-e_a-e_b), t1 ~ N(0, t1_sigma^2). */
int xt_synth_upload(xt_ctx* ctx, uint64_t seed, double sigma, double t1_sigma);
/* Same generator, writing host arrays (any output may be NULL to skip it). */
int xt_synth_closed_shell(int O, int V, uint64_t seed, double sigma, double t1_sigma,
                          double* eps_occ, double* eps_vir, double* t1, double* t2, double* vvoo,
                          double* ovoo, double* vvvo);

typedef struct xt_opts {
  int   precision;   /* XT_PREC_FP64 or XT_PREC_MIXED                                */
  int   schedule;    /* XT_SCHED_*                                                   */
  int   use_graph;   /* 1: capture the launch sequence in a hipGraph and replay it   */
  int   record_task_time; /* 1: record per-task device timestamps                    */
  int   workgroups;  /* persistent grid size (0 = library default)                   */
  int   progress_ms; /* > 0: while the kernel runs, print the fraction of work-queue
                        tickets taken to stderr every progress_ms milliseconds       */
  int   reserved[2];
} xt_opts;

/* Default options (FP64, driver order, graph on, timing off). */
void xt_default_opts(xt_opts* o);

/* Builds the device operand form a precision needs (for XT_PREC_MIXED the
 * row-scaled FP16 hi/lo split of every contraction operand), which xt_run
 * otherwise builds on its first call at that precision.  No kernel of the
 * energy itself runs. */
int xt_prepare(xt_ctx* ctx, int precision);

/*
 * Runs every task.  e45[0] = E[4] ([T] correction), e45[1] = E[5] ((T)
 * correction, which includes E[4]), summed in task order in FP64 with the
 * factors applied per task.  task_partials[2*n] (optional) receives each
 * task's (E[4], E[5]) contribution, factor included; task_ms[n] (optional)
 * receives per-task device time in ms when opts->record_task_time is set.
 * XT_ERANGE if an energy partial was not finite or beyond 2^62 Eh (it cannot
 * enter the exact fixed-point sums); e45 and task_partials are then NaN.
 */
int xt_run(xt_ctx* ctx, const xt_opts* opts, double* e45, double* task_partials, float* task_ms);

/* Runs only the listed tasks (indices into the task list, no repeats).
 * e45 = their summed (E[4], E[5]) in the listed order; task_partials[2*n] and
 * task_ms[n] (optional) follow the listed order. */
int xt_run_tasks(xt_ctx* ctx, const xt_opts* opts, long n, const long* which, double* e45,
                 double* task_partials, float* task_ms);

/* Device memory the inputs of the current tiling need, in bytes: bytes[0]
 * the resident FP64 operand arrays, bytes[1] what an upload holds beyond them
 * until the layout is built (nchol = 0: xt_upload_closed_shell; nchol > 0:
 * xt_upload_cholesky with that many Cholesky vectors), bytes[2] the FP16
 * split copies the first mixed-precision run adds.  Host only.  The uploads
 * and the first mixed run check these figures against the device's free
 * memory before allocating and fail with XT_ENOMEM and the figures
 * otherwise; replaces check_memory_req (hybrid.cpp:19-41). */
int xt_memory_required(xt_ctx* ctx, long nchol, double* bytes);

/* Largest number of device bytes the last upload (xt_upload_closed_shell,
 * xt_synth_upload or xt_upload_cholesky) held at once through its own
 * allocations: resident operands plus the staging still alive (workspace
 * rocBLAS allocates for itself is not counted); 0 before the first upload.
 * Host only.  For checking xt_memory_required's operands + staging figure,
 * which the uploads test against free memory before they allocate. */
int xt_upload_peak_bytes(xt_ctx* ctx, double* bytes);

/* E[4], E[5] from per-task contributions: task_partials[2*n] in task-list
 * order, factor included, as xt_run returns them.  The sum follows exactly
 * the fixed order of xt_run's device reduction.  A task's contribution is an
 * exact fixed-point sum and does not depend on which other tasks share its
 * run.  So a run that stopped part-way, with its finished tasks' partials
 * kept and the rest computed by xt_run_tasks, gives the bit-identical totals
 * of one full xt_run.  Host only; n = 0 gives 0.  No counterpart in the
 * reference, whose (T) restart re-reads T1/T2/V2 from disk and recomputes
 * every task (ccsd_t.cpp:162-201). */
int xt_sum_task_energies(long n, const double* task_partials, double* e45);

/* Slot codes of one task (introspection for tests and tools): codes[0] d1,
 * codes[1] d2 (6 bits per output bipartition g = 3*hs+ps: enabled, negative,
 * a1, a2, b1, b2), codes[2] s1 (4 bits per slot 3*pa+hi: enabled, negative,
 * b1, b2). */
int xt_task_codes(xt_ctx* ctx, long task, uint64_t* codes);

/* Work-queue layout a run over `which` (n tasks; n = 0: every task) would
 * launch with under opts' precision and schedule (introspection for tests
 * and tools; host only, no device work).  Eight queues, one per XCD: queue q
 * holds the entries qbeg[q]..qbeg[q+1]-1 (qbeg[9]); entry e is task qtask[e]
 * (n or ntasks entries) owning the queue's tickets from qtick[e]; qtotal[8]
 * tickets per queue.  nbox[6*ntasks]: every task's box grid at that
 * precision; glog[6]: the ticket group shape in use.  Replaces the NXTVAL
 * task counter of ccsd_t_fused_driver.hpp:169-172,381,456. */
int xt_queue_layout(xt_ctx* ctx, const xt_opts* opts, long n, const long* which, int* qbeg,
                    int* qtask, long long* qtick, long long* qtotal, int* nbox, int* glog);

/* ------------------------------------ device-side gather (per-task path) */
/*
 * Replaces the host gathers ccsd_t_data_s1_new (ccsd_t_all_fused_singles.hpp:559-580),
 * ccsd_t_data_d1_new (ccsd_t_all_fused_doubles1.hpp:491-514) and
 * ccsd_t_data_d2_new (ccsd_t_all_fused_doubles2.hpp:547-570), called per task
 * from ccsd_t_all_fused.hpp:149-178.  Instead of TAMM blocks (Tensor::get)
 * permuted on the host by hptt, the packed buffers of one task are written
 * on `stream` from the context's device-resident inputs (xt_upload_*), so
 * they feed xt_fully_fused_ccsd_t_gpu without a host round trip.
 *
 * task = (t_h1b, t_h2b, t_h3b, t_p4b, t_p5b, t_p6b), spin-orbital tile
 * indices of the context's tiling (occupied tiles 0..noab-1, virtual
 * noab..noab+nvab-1).  The six device buffers are sized by xt_packed_sizes;
 * only the blocks of enabled slots are written (slot k of a buffer starts at
 * k * size / max_slots), as in the reference.  The host arrays are filled as
 * the reference fills df_simple_*_size / df_simple_*_exec / df_num_*_enabled:
 * s1_size[6], s1_exec[9], d1_size[7*noab], d1_exec[9*noab], d2_size[7*nvab],
 * d2_exec[9*nvab] (exec -1 where no slot runs), counts[3] = enabled s1, d1, d2
 * slots.  The context's block list is reused by the next call: issue calls
 * of one context on one stream.
 */
int xt_gather_task(xt_ctx* ctx, const long task[6], void* stream, double* d1_t2, double* d1_v2,
                   double* d2_t2, double* d2_v2, double* s1_t1, double* s1_v2, int* s1_size,
                   int* s1_exec, int* d1_size, int* d1_exec, int* d2_size, int* d2_exec,
                   int counts[3]);
/* Buffer sizes of the packed path, the driver's size_T_s1_t1, size_T_s1_v2,
 * size_T_d1_t2, size_T_d1_v2, size_T_d2_t2, size_T_d2_v2 (doubles) and
 * max_d1_kernels_pertask, max_d2_kernels_pertask (ccsd_t_fused_driver.hpp:176-195). */
int xt_packed_sizes(xt_ctx* ctx, long sizes[8]);
/* Dense spin-orbital amplitudes from the context's closed-shell ones, on
 * `stream`: the expansion of cd_ccsd_cs_ann.cpp:932-976 (t1_aa into both spin
 * blocks of d_t1; t2_aaaa = t2_bbbb = t_ij^ab - t_ij^ba, t2_abab = t2_baba =
 * t_ij^ab, t2_abba = t2_baab = -t_ij^ab into d_t2), with spin-orbital index
 * order [alpha, beta] per space.  t1_so is [2V][2O] and t2_so is
 * [2V][2V][2O][2O], device memory; either may be NULL. */
int xt_expand_amplitudes(xt_ctx* ctx, void* stream, double* t1_so, double* t2_so);
/* Doubles written by the context's last xt_gather_task (the enabled blocks). */
int xt_gather_last_elems(xt_ctx* ctx, long* n);

/* Device time of the last xt_run's main kernel (ms), from HIP events. */
int xt_last_kernel_ms(xt_ctx* ctx, double* ms);

/* ------------------------------------------------------------------------
 * Tile-shape autotuner (no counterpart in the reference, whose GPU kernel
 * runs one 4^6 block per workgroup in a fixed order, nontc:2801-2919).
 *
 * Each kernel works on boxes of t3 (4^6 sub-tiles in FP64, (4,4,4|8,8,8)
 * boxes in mixed precision); the work queue hands them out in groups of 64
 * neighbouring boxes, 2^glog[p] along position p of (h1,h2,h3,p4,p5,p6).
 * The group shape decides which operand panels the workgroups of one XCD
 * share in its L2.  It never changes the energies (bit for bit).
 * ------------------------------------------------------------------------ */
#define XT_TUNE_CANDIDATES 6

/* glog[6]: non-negative, summing to 6.  Default {1,1,1,1,1,1}. */
int xt_set_group_shape(xt_ctx* ctx, int precision, const int* glog);
int xt_get_group_shape(xt_ctx* ctx, int precision, int* glog);
/* Times every candidate group shape (XT_TUNE_CANDIDATES of them) on
 * n_sample tasks spread over the flop range at opts->precision, installs the
 * fastest for that precision and returns its index in *best; times_ms
 * (optional, XT_TUNE_CANDIDATES entries) receives each candidate's kernel
 * time and shapes (optional, 6 * XT_TUNE_CANDIDATES) the candidates. */
int xt_autotune(xt_ctx* ctx, const xt_opts* opts, long n_sample, int* best, double* times_ms,
                int* shapes);

/* ------------------------------------------------------------------------
 * Learned task-cost model for the work-queue scheduler (no counterpart in
 * the reference, whose tasks are handed out in driver order by NXTVAL,
 * ccsd_t_fused_driver.hpp:381).
 *
 * A 7 -> 64 -> 64 -> 1 FP32 perceptron (ReLU hidden layers) whose training
 * (forward, back-propagation, Adam) and inference run as HIP kernels on the
 * model's device.  Inputs per task (xt_task_features): the six tile extents
 * / 48 and the precision flag (0 FP64, 1 mixed).  The scheduler uses it to
 * predict log(task time / analytic flops); a task's cost is its flop count
 * times the predicted time per flop.
 * ------------------------------------------------------------------------ */
#define XT_MLP_IN 7
#define XT_MLP_HIDDEN 64
#define XT_MLP_PARAMS 4737 /* W1[7][64] b1[64] W2[64][64] b2[64] w3[64] b3 */

typedef struct xt_mlp xt_mlp;

/* He-initialised parameters from a seeded generator (biases zero). */
int xt_mlp_create(int device, uint64_t seed, xt_mlp** out);
int xt_mlp_destroy(xt_mlp* m);
/* Parameters in the layout above; set also clears the Adam state. */
int xt_mlp_get_params(xt_mlp* m, float* params);
int xt_mlp_set_params(xt_mlp* m, const float* params);
/* Mean squared error over the batch x[n][7] -> t[n] and its gradient with
 * respect to every parameter (grad[XT_MLP_PARAMS]); parameters unchanged. */
int xt_mlp_loss_grad(xt_mlp* m, long n, const float* x, const float* t, float* loss, float* grad);
/* `steps` full-batch Adam steps (beta 0.9/0.999, eps 1e-8) at rate lr,
 * replayed from one captured graph; loss_hist[steps] (optional) receives the
 * loss before each step.  Deterministic. */
int xt_mlp_train(xt_mlp* m, long n, const float* x, const float* t, int steps, float lr,
                 float* loss_hist);
/* y[n] = model(x[n][7]). */
int xt_mlp_predict(xt_mlp* m, long n, const float* x, float* y);

/* Model inputs of every task of the tiling at a precision: feats[n_tasks][7]. */
int xt_task_features(xt_ctx* ctx, int precision, float* feats);
/* Per-task costs used by XT_SCHED_LEARNED (any positive unit). */
int xt_set_task_cost(xt_ctx* ctx, const double* cost);

typedef struct xt_learn_stats {
  long   n_sample;         /* tasks timed per precision                        */
  double final_loss;       /* training loss (standardised log time per flop)   */
  double fit_rmse_model;   /* RMS error of log(time) on the timed tasks: model  */
  double fit_rmse_flops;   /* same for time = flops / (best single rate)       */
  double calib_ms;         /* wall time of the timing runs                      */
  double train_ms;         /* wall time of training and prediction              */
} xt_learn_stats;

/* Learned schedule: times n_sample tasks (spread over the flop range) at
 * each precision in precision_mask (bit 0 FP64, bit 1 mixed), trains the
 * model on log(time/flops) (standardised), predicts every task at
 * opts->precision and installs flops * predicted time per flop as the
 * XT_SCHED_LEARNED costs.  The analytic flop count stays the fallback: a
 * non-finite or non-positive prediction leaves that task at its flop cost. */
int xt_learn_schedule(xt_ctx* ctx, xt_mlp* m, const xt_opts* opts, long n_sample,
                      int precision_mask, int steps, float lr, xt_learn_stats* stats);

#ifdef __cplusplus
}
#endif
#endif /* EXACHEM_T_H */
