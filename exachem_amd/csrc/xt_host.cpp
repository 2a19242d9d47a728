// Host side of the MI355X (T) path: tiling and task planning (the fused
// driver's task loop), the seeded synthetic generator, HBM residency of the
// inputs, the persistent launch and the C ABI of include/exachem_t.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/exachem_t.h"
#include "xt_device.h"
#include "xt_internal.h"
#include "xt_queue.h"

namespace {
thread_local std::string g_err;
}  // namespace

namespace xt {
int set_error(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
}  // namespace xt

namespace {

inline int fail(int code, const std::string& msg) { return xt::set_error(code, msg); }

#define XT_HIP(call)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if(e_ != hipSuccess)                                                                \
      return fail(XT_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));          \
  } while(0)

// ------------------------------------------------------------- generator
inline uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}
inline uint64_t hash3(uint64_t seed, uint64_t stream, uint64_t idx) {
  return mix64(mix64(seed ^ (stream * 0xD6E8FEB86659FD93ULL)) ^ mix64(idx));
}
inline double u01(uint64_t h) { return ((double) (h >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
inline double gauss(uint64_t seed, uint64_t stream, uint64_t idx) {
  const double u1 = u01(hash3(seed, stream, 2 * idx));
  const double u2 = u01(hash3(seed, stream, 2 * idx + 1));
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586476925286766559 * u2);
}

template<class F>
void parallel_for(long n, F f) {
  unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  if(n < 64) nt = 1;
  std::vector<std::thread> th;
  std::atomic<long>        next{0};
  auto                     body = [&]() {
    for(;;) {
      long i = next.fetch_add(1);
      if(i >= n) break;
      f(i);
    }
  };
  for(unsigned t = 1; t < nt; t++) th.emplace_back(body);
  body();
  for(auto& t: th) t.join();
}

enum Stream : uint64_t { S_EO = 1, S_EV = 2, S_VVOO = 3, S_OVOO = 4, S_VVVO = 5, S_T1 = 6 };

int synth(int O, int V, uint64_t seed, double sigma, double t1s, double* eo, double* ev,
          double* t1, double* t2, double* vvoo, double* ovoo, double* vvvo) {
  if(O <= 0 || V <= 0) return fail(XT_EINVAL, "synth: O and V must be positive");
  std::vector<double> eo_(O), ev_(V);
  for(int i = 0; i < O; i++) eo_[i] = -2.0 + 1.7 * u01(hash3(seed, S_EO, i));
  for(int a = 0; a < V; a++) ev_[a] = 0.1 + 2.9 * u01(hash3(seed, S_EV, a));
  std::sort(eo_.begin(), eo_.end());
  std::sort(ev_.begin(), ev_.end());
  if(eo) std::copy(eo_.begin(), eo_.end(), eo);
  if(ev) std::copy(ev_.begin(), ev_.end(), ev);
  const long o = O, v = V;
  if(t1)
    for(long a = 0; a < v; a++)
      for(long i = 0; i < o; i++) t1[a * o + i] = t1s * gauss(seed, S_T1, a * o + i);
  if(vvoo || t2) {
    // (ai|bj) = (bj|ai): value drawn at the lexicographically smaller of (a,i,b,j), (b,j,a,i)
    parallel_for(v, [&](long a) {
      for(long b = 0; b < v; b++)
        for(long i = 0; i < o; i++)
          for(long j = 0; j < o; j++) {
            bool   sw = (b < a) || (b == a && j < i);
            long   a_ = sw ? b : a, i_ = sw ? j : i, b_ = sw ? a : b, j_ = sw ? i : j;
            double x  = sigma * gauss(seed, S_VVOO, ((a_ * o + i_) * v + b_) * o + j_);
            size_t id = ((a * v + b) * o + i) * o + j;
            if(vvoo) vvoo[id] = x;
            if(t2) t2[id] = x / ((eo_[i] + eo_[j]) - (ev_[a] + ev_[b]));
          }
    });
  }
  if(ovoo) {
    parallel_for(o, [&](long i) {
      for(long a = 0; a < v; a++)
        for(long j = 0; j < o; j++)
          for(long k = 0; k < o; k++) {
            long jm = std::min(j, k), kM = std::max(j, k);
            ovoo[((i * v + a) * o + j) * o + k] =
              sigma * gauss(seed, S_OVOO, ((i * v + a) * o + jm) * o + kM);
          }
    });
  }
  if(vvvo) {
    parallel_for(v, [&](long a) {
      for(long b = 0; b < v; b++) {
        long am = std::min(a, b), bM = std::max(a, b);
        for(long c = 0; c < v; c++)
          for(long i = 0; i < o; i++)
            vvvo[((a * v + b) * v + c) * o + i] =
              sigma * gauss(seed, S_VVVO, ((am * v + bM) * v + c) * o + i);
      }
    });
  }
  return XT_OK;
}

// ------------------------------------------------------------- planning
inline void pair_of(int s, int& a, int& b) {
  a = (s == 0) ? 1 : 0;
  b = (s == 2) ? 1 : 2;
}
inline int sgn_single(int x) { return x == 1 ? -1 : 1; }

struct Plan {
  int                     O = 0, V = 0;
  bool                    restricted = true;
  std::vector<int>        k_range, k_spin, spatial_off;
  int                     noab = 0, nvab = 0;
  std::vector<long>       tasks;  // 6 per task, tile ids
  std::vector<double>     factors;
  std::vector<xt::TaskDev> dev;
};

// Slot logic of one task, derived from the tile spins.  It restates the
// enable and operand rules of the gathers (singles.hpp:100-119,
// doubles1.hpp:94-117, doubles2.hpp:98-121): a d1 equation with single hole
// hs and particle pair (pA,pB) runs over h7 of spin s(pA)+s(pB)-s(hs), a d2
// equation with hole pair (hA,hB) and single particle ps over p7 of spin
// s(hA)+s(hB)-s(ps); s1 needs s(pa) = s(hi).  The other enable tests of the
// gathers (tile order, spin total, restricted total != 12) always hold for
// tasks that pass the driver filter, since every tuple is a permutation of
// the task.
void slot_codes(const int* sp, uint64_t& d1, uint64_t& d2, uint64_t& s1, int& nd1, int& nd2,
                int& ns1) {
  d1 = d2 = s1 = 0;
  nd1 = nd2 = ns1 = 0;
  for(int hs = 0; hs < 3; hs++)
    for(int ps = 0; ps < 3; ps++) {
      const int g = 3 * hs + ps;
      int       hA, hB, pA, pB;
      pair_of(hs, hA, hB);
      pair_of(ps, pA, pB);
      const int sHS = sp[hs], sHA = sp[hA], sHB = sp[hB];
      const int sPS = sp[3 + ps], sPA = sp[3 + pA], sPB = sp[3 + pB];
      const bool neg = (-sgn_single(hs) * sgn_single(ps)) < 0;
      {
        const int s7 = sPA + sPB - sHS;
        if(s7 == 1 || s7 == 2) {
          uint64_t c = 1 | (neg ? 2 : 0);
          if(sPA == sHS && sPB == s7) c |= 4;   // t(pA,pB,hs,h7) term
          if(sPA == s7 && sPB == sHS) c |= 8;   // -t(pA,pB,h7,hs) term
          if(sHA == s7 && sHB == sPS) c |= 16;  // (hA h7|hB ps)
          if(sHB == s7 && sHA == sPS) c |= 32;  // -(hB h7|hA ps)
          d1 |= c << (6 * g);
          nd1++;
        }
      }
      {
        const int s7 = sHA + sHB - sPS;
        if(s7 == 1 || s7 == 2) {
          uint64_t c = 1 | (neg ? 2 : 0);
          if(sHS == sPA && s7 == sPB) c |= 4;   // (hs pA|p7 pB)
          if(sHS == sPB && s7 == sPA) c |= 8;   // -(hs pB|p7 pA)
          if(sPS == sHA && s7 == sHB) c |= 16;  // t(ps,p7,hA,hB)
          if(sPS == sHB && s7 == sHA) c |= 32;  // -t(p7,ps,hA,hB)
          d2 |= c << (6 * g);
          nd2++;
        }
      }
    }
  for(int pa = 0; pa < 3; pa++)
    for(int hi = 0; hi < 3; hi++) {
      const int slot = 3 * pa + hi;
      if(sp[3 + pa] != sp[hi]) continue;
      int hj, hk, pb, pc;
      pair_of(hi, hj, hk);
      pair_of(pa, pb, pc);
      uint64_t c = 1 | ((sgn_single(hi) * sgn_single(pa)) < 0 ? 2 : 0);
      if(sp[hk] == sp[3 + pc] && sp[hj] == sp[3 + pb]) c |= 4;  // (hk pc|hj pb)
      if(sp[hk] == sp[3 + pb] && sp[hj] == sp[3 + pc]) c |= 8;  // -(hk pb|hj pc)
      s1 |= c << (4 * slot);
      ns1++;
    }
}

int build_plan(Plan& P, int O, int V, const std::vector<int>& occ, const std::vector<int>& vir,
               bool restricted) {
  P = Plan();
  P.O = O;
  P.V = V;
  P.restricted = restricted;
  // |occ_a|occ_b|virt_a|virt_b| (cholesky_2e.cpp:192-277), k_spin (ccsd_t.cpp:245-249)
  int off = 0;
  for(int half = 0; half < 2; half++)
    for(int t: occ) {
      P.k_range.push_back(t);
      P.k_spin.push_back(half + 1);
      P.spatial_off.push_back(off % O);
      off += t;
    }
  off = 0;
  for(int half = 0; half < 2; half++)
    for(int t: vir) {
      P.k_range.push_back(t);
      P.k_spin.push_back(half + 1);
      P.spatial_off.push_back(off % V);
      off += t;
    }
  P.noab = 2 * (int) occ.size();
  P.nvab = 2 * (int) vir.size();
  const int noab = P.noab, nvab = P.nvab;
  const auto& ks = P.k_spin;
  // task loop of the fused driver, seq-h3b variant (ccsd_t_fused_driver.hpp:368-395)
  for(int h1 = 0; h1 < noab; h1++)
    for(int p4 = noab; p4 < noab + nvab; p4++)
      for(int h2 = h1; h2 < noab; h2++)
        for(int p5 = p4; p5 < noab + nvab; p5++)
          for(int p6 = p5; p6 < noab + nvab; p6++)
            for(int h3 = h2; h3 < noab; h3++) {
              if(ks[p4] + ks[p5] + ks[p6] != ks[h1] + ks[h2] + ks[h3]) continue;
              if(restricted && ks[p4] + ks[p5] + ks[p6] + ks[h1] + ks[h2] + ks[h3] > 8) continue;
              double f = restricted ? 2.0 : 1.0;
              if(p4 == p5 && p5 == p6) f /= 6.0;
              else if(p4 == p5 || p5 == p6) f /= 2.0;
              if(h1 == h2 && h2 == h3) f /= 6.0;
              else if(h1 == h2 || h2 == h3) f /= 2.0;
              const long t[6] = {h1, h2, h3, p4, p5, p6};
              P.tasks.insert(P.tasks.end(), t, t + 6);
              P.factors.push_back(f);
            }
  const long nt = (long) P.factors.size();
  P.dev.resize(nt);
  for(long k = 0; k < nt; k++) {
    xt::TaskDev& T = P.dev[k];
    std::memset(&T, 0, sizeof(T));
    int sp[6];
    T.n_sub = 1;
    for(int p = 0; p < 6; p++) {
      const long b = P.tasks[6 * k + p];
      T.dim[p]     = P.k_range[b];
      T.off[p]     = P.spatial_off[b];
      T.tile[p]    = p < 3 ? (int) (b % (noab / 2)) : (int) ((b - noab) % (nvab / 2));
      T.nsub[p]    = (T.dim[p] + 3) / 4;
      T.n_sub *= T.nsub[p];
      sp[p] = P.k_spin[b];
    }
    int nd1, nd2, ns1;
    slot_codes(sp, T.d1code, T.d2code, T.s1code, nd1, nd2, ns1);
    T.factor     = P.factors[k];
    for(int p = 0; p < 6; p++) T.nsubm[p] = (T.dim[p] + (p < 3 ? 3 : 7)) / (p < 3 ? 4 : 8);
    double prod = 1;
    for(int p = 0; p < 6; p++) prod *= T.dim[p];
    // helper_calculate_num_ops (fused_common.hpp:131-261): 2*prod per s1 slot,
    // 2*prod*|h7 range| per d1 slot, 2*prod*|p7 range| per d2 slot; closed-shell
    // spin halves have O (V) orbitals each.
    T.flops = 2.0 * prod * (ns1 + (double) nd1 * O + (double) nd2 * V);
  }
  return XT_OK;
}

int current_device_is_gfx950(int dev) {
  hipDeviceProp_t prop;
  if(hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
  return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
}

}  // namespace

// =============================================================== context
struct xt_ctx {
  int         device = 0;
  hipStream_t stream = nullptr;
  // host-mapped mirror of the queue ticket counters: the kernel stores the
  // tickets it takes, the host reads them while the persistent kernel runs
  unsigned long long* h_prog = nullptr;
  unsigned long long* d_prog = nullptr;
  int         n_cu   = 0;
  Plan        plan;
  bool        have_plan = false, have_inputs = false;
  // device inputs (kernel layouts)
  xt::SpatialBufs in    = {};
  xt::SplitBufs   split = {};  // FP16 split operands (built on the first mixed run)
  bool            have_split = false;
  // device plan and run state
  xt::TaskDev*        d_tasks  = nullptr;
  int*                d_qtask  = nullptr;
  int*                d_qbeg   = nullptr;
  long long*          d_qtick  = nullptr;
  long long*          d_qtotal = nullptr;
  unsigned long long* d_qctr   = nullptr;
  unsigned long long* d_acc    = nullptr;
  unsigned long long* d_cost   = nullptr;
  double *            d_task_e = nullptr, *d_e45 = nullptr;
  long                dev_tasks_n = -1;
  // device-side gather: block descriptors of the last call (device copy and
  // the host list it was uploaded from, kept alive until the next call)
  xt::GatherBlock*              d_gather     = nullptr;
  long                          d_gather_cap = 0;
  std::vector<xt::GatherBlock>  h_gather;
  long                          gather_elems = 0;  // doubles written by the last gather
  hipEvent_t          ev0 = nullptr, ev1 = nullptr;
  double              last_ms = 0.0;
  double              upload_peak = 0.0;  // largest device bytes the last upload held at once
  std::vector<double> learned_cost;  // per-task predicted cost (learned schedule)
  // ticket group shapes (log2 extents per position), set by the autotuner
  int glog_fp64[6]  = {1, 1, 1, 1, 1, 1};
  int glog_mixed[6] = {1, 1, 1, 1, 1, 1};
  // Captured launch sequence (zero fills, persistent kernel, reductions).  It
  // reads the queues from the device buffers above, so one instantiation
  // serves every schedule and task subset of the same precision / grid /
  // timing mode; it is dropped whenever a buffer it names is reallocated.
  hipGraphExec_t graph = nullptr;
  struct GraphKey {
    int         precision = -1, grid = 0, timed = 0;
    long        n_tasks   = -1;
    const void* inputs    = nullptr;
    bool        operator==(const GraphKey& k) const {
      return precision == k.precision && grid == k.grid && timed == k.timed &&
             n_tasks == k.n_tasks && inputs == k.inputs;
    }
  } graph_key;

  void drop_graph() {
    if(graph) (void) hipGraphExecDestroy(graph);
    graph     = nullptr;
    graph_key = GraphKey();
  }

  void free_inputs() {
    drop_graph();
    double** ps[] = {&in.eo,  &in.ev,  &in.t1,  &in.t2h, &in.t2ha, &in.ovk,
                     &in.ovka, &in.xv, &in.xva, &in.t2p, &in.t2pa, &in.vvoo};
    for(double** p: ps)
      if(*p) {
        (void) hipFree(*p);
        *p = nullptr;
      }
    have_inputs = false;
    free_split();
  }
  void free_split() {
    drop_graph();
    for(int j = 0; j < xt::kOperands; j++) {
      if(split.v[j]) (void) hipFree(split.v[j]);
      if(split.eblk[j]) (void) hipFree(split.eblk[j]);
      split.v[j] = nullptr;
      split.eblk[j] = nullptr;
    }
    if(split.zero) (void) hipFree(split.zero);
    split.zero = nullptr;
    if(split.pscale) (void) hipFree(split.pscale);
    split.pscale   = nullptr;
    split.pscale_n = 0;
    have_split = false;
  }
  void free_plan_dev() {
    drop_graph();
    void* ps[] = {d_tasks, d_qtask, d_qbeg, d_qtick, d_qtotal, d_qctr, d_acc, d_cost, d_task_e, d_e45};
    for(void* p: ps)
      if(p) (void) hipFree(p);
    d_tasks  = nullptr;
    d_qtask  = nullptr;
    d_qbeg   = nullptr;
    d_qtick  = nullptr;
    d_qtotal = nullptr;
    d_qctr   = nullptr;
    d_acc    = nullptr;
    d_cost   = nullptr;
    d_task_e = d_e45 = nullptr;
    dev_tasks_n      = -1;
  }
};

extern "C" {

const char* xt_last_error(void) { return g_err.c_str(); }
const char* xt_version(void) { return "exachem_amd-xt 0.1 (gfx950)"; }

int xt_device_available(void) {
  int n = 0;
  if(hipGetDeviceCount(&n) != hipSuccess || n <= 0) return 0;
  return current_device_is_gfx950(0);
}

int xt_create(int device, xt_ctx** out) {
  if(!out) return fail(XT_EINVAL, "xt_create: out is null");
  *out = nullptr;
  if(device == -1) {  // host-only context: planning (tiling, tasks, flops), no HIP calls
    auto* c   = new xt_ctx();
    c->device = -1;
    *out      = c;
    return XT_OK;
  }
  int n = 0;
  if(hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return fail(XT_ENODEV, "xt_create: no HIP device");
  if(device < 0 || device >= n) return fail(XT_EINVAL, "xt_create: bad device index");
  if(!current_device_is_gfx950(device))
    return fail(XT_ENODEV, "xt_create: device is not gfx950");
  auto* c   = new xt_ctx();
  c->device = device;
  XT_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  XT_HIP(hipGetDeviceProperties(&prop, device));
  c->n_cu = prop.multiProcessorCount;
  XT_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  XT_HIP(hipHostMalloc((void**) &c->h_prog, sizeof(unsigned long long) * xt::kQueues,
                       hipHostMallocMapped | hipHostMallocCoherent));
  XT_HIP(hipHostGetDevicePointer((void**) &c->d_prog, c->h_prog, 0));
  XT_HIP(hipEventCreate(&c->ev0));
  XT_HIP(hipEventCreate(&c->ev1));
  *out = c;
  return XT_OK;
}

int xt_destroy(xt_ctx* c) {
  if(!c) return XT_OK;
  if(c->device < 0) {
    delete c;
    return XT_OK;
  }
  (void) hipSetDevice(c->device);
  (void) hipStreamSynchronize(c->stream);
  c->free_inputs();
  c->free_plan_dev();
  if(c->ev0) (void) hipEventDestroy(c->ev0);
  if(c->ev1) (void) hipEventDestroy(c->ev1);
  if(c->stream) (void) hipStreamDestroy(c->stream);
  if(c->h_prog) (void) hipHostFree(c->h_prog);
  if(c->d_gather) (void) hipFree(c->d_gather);
  delete c;
  return XT_OK;
}

int xt_set_tiling(xt_ctx* c, int O, int V, int n_occ_tiles, const int* occ_tiles,
                  int n_vir_tiles, const int* vir_tiles, int is_restricted) {
  if(!c || !occ_tiles || !vir_tiles || n_occ_tiles <= 0 || n_vir_tiles <= 0)
    return fail(XT_EINVAL, "xt_set_tiling: bad arguments");
  std::vector<int> occ(occ_tiles, occ_tiles + n_occ_tiles), vir(vir_tiles, vir_tiles + n_vir_tiles);
  long so = 0, sv = 0;
  for(int t: occ) {
    if(t <= 0) return fail(XT_EINVAL, "xt_set_tiling: empty occupied tile");
    so += t;
  }
  for(int t: vir) {
    if(t <= 0) return fail(XT_EINVAL, "xt_set_tiling: empty virtual tile");
    sv += t;
  }
  if(so != O || sv != V) return fail(XT_EINVAL, "xt_set_tiling: tiles do not sum to O / V");
  if(c->have_inputs && (c->plan.O != O || c->plan.V != V)) c->free_inputs();
  c->free_split();  // block exponents follow the tiling
  int rc = build_plan(c->plan, O, V, occ, vir, is_restricted != 0);
  if(rc) return rc;
  c->have_plan = true;
  c->free_plan_dev();
  return XT_OK;
}

int xt_num_tasks(xt_ctx* c, long* n) {
  if(!c || !n) return fail(XT_EINVAL, "xt_num_tasks: null argument");
  if(!c->have_plan) return fail(XT_ESTATE, "xt_num_tasks: no tiling");
  *n = (long) c->plan.factors.size();
  return XT_OK;
}

int xt_task_list(xt_ctx* c, long* tasks, double* factors) {
  if(!c) return fail(XT_EINVAL, "xt_task_list: null context");
  if(!c->have_plan) return fail(XT_ESTATE, "xt_task_list: no tiling");
  if(tasks) std::copy(c->plan.tasks.begin(), c->plan.tasks.end(), tasks);
  if(factors) std::copy(c->plan.factors.begin(), c->plan.factors.end(), factors);
  return XT_OK;
}

int xt_flops(xt_ctx* c, double* per_task, double* total, double* parts) {
  if(!c) return fail(XT_EINVAL, "xt_flops: null context");
  if(!c->have_plan) return fail(XT_ESTATE, "xt_flops: no tiling");
  long double tot = 0, s1 = 0, d1 = 0, d2 = 0;
  for(size_t k = 0; k < c->plan.dev.size(); k++) {
    const auto& T = c->plan.dev[k];
    if(per_task) per_task[k] = T.flops;
    tot += T.flops;
    long double prod = 1;
    for(int p = 0; p < 6; p++) prod *= T.dim[p];
    int nd1 = 0, nd2 = 0, ns1 = 0;
    for(int g = 0; g < 9; g++) {
      nd1 += (T.d1code >> (6 * g)) & 1;
      nd2 += (T.d2code >> (6 * g)) & 1;
      ns1 += (T.s1code >> (4 * g)) & 1;
    }
    s1 += 2 * prod * ns1;
    d1 += 2 * prod * nd1 * c->plan.O;
    d2 += 2 * prod * nd2 * c->plan.V;
  }
  if(total) *total = (double) tot;
  if(parts) {
    parts[0] = (double) s1;
    parts[1] = (double) d1;
    parts[2] = (double) d2;
  }
  return XT_OK;
}

int xt_synth_closed_shell(int O, int V, uint64_t seed, double sigma, double t1_sigma,
                          double* eps_occ, double* eps_vir, double* t1, double* t2, double* vvoo,
                          double* ovoo, double* vvvo) {
  return synth(O, V, seed, sigma, t1_sigma, eps_occ, eps_vir, t1, t2, vvoo, ovoo, vvvo);
}

namespace {
// Device bytes of the closed-shell inputs of a tiling, following the
// allocations of build_operands, build_xva and build_split: the resident FP64
// operands, what an upload holds beyond them until the layout is built
// (nchol = 0: the closed-shell upload; nchol > 0: the Cholesky upload, whose
// high point is the largest of its GEMM, layout and xva steps), and the FP16
// split copies of the mixed form.  Per-task arrays (bytes per task) are left
// out.
struct MemReq {
  double operands, staging, split;
};
MemReq mem_req(size_t O, size_t V, long nchol) {
  const double o = (double) O, v = (double) V;
  const double op = (double) ((O + 7) / 8 * 8), vp = (double) ((V + 7) / 8 * 8);
  const double ko = (double) ((O + 15) / 16 * 16), kv = (double) ((V + 15) / 16 * 16);
  MemReq       r;
  r.operands = 8 * (2 * v * v * o * op + 2 * o * o * v * op + 2 * v * v * o * vp + 2 * o * o * v * vp +
                    v * v * o * o + v * o + o + v);
  r.split    = 4 * (2 * v * v * o * ko + 2 * o * o * v * ko + 2 * v * v * o * kv + 2 * o * o * v * kv) +
            2 * std::max(v * v * o, o * o * v);
  const double t2 = 8 * v * v * o * o, ov = 8 * o * o * o * v, xv = 8 * v * v * o * vp;
  if(nchol <= 0) {
    // t2, ovoo and vvvo as uploaded; vvvo is freed once xv exists, before xva
    // (at least as large) is allocated, so the peak holds t2 and ovoo beside
    // all the operands
    r.staging = t2 + ov + std::max(0.0, 8 * v * v * v * o - xv);
  }
  else {
    const double nmo = o + v, k = (double) nchol;
    const double gemm = 8 * (v * v * o * o + v * o + o + v) + xv + t2 + 8 * nmo * nmo * k + ov +
                        8 * (2 * v * o * k + o * o * k + v * v * o * o);
    const double layout = r.operands - xv + t2 + ov + 8 * v * v * o * o;
    const double xva    = r.operands + t2 + ov;
    r.staging           = std::max({gemm, layout, xva}) - r.operands;
  }
  return r;
}

// Fails with XT_ENOMEM before anything is allocated when a + b bytes (named
// na, nb) exceed the device's free memory (hybrid.cpp:19-41 checks the (T)
// requirement against the device's total memory the same way).
int check_free(const char* what, double a, const char* na, double b, const char* nb) {
  size_t fr = 0, tot = 0;
  if(hipMemGetInfo(&fr, &tot) != hipSuccess) return XT_OK;  // no figure: hipMalloc reports
  if(a + b <= (double) fr) return XT_OK;
  const double gib = 1024.0 * 1024.0 * 1024.0;
  char         buf[384];
  snprintf(buf, sizeof buf,
           "%s: GPU memory not sufficient for this tiling: %.1f GiB needed (%.1f GiB %s + %.1f "
           "GiB %s), %.1f GiB free of %.1f GiB",
           what, (a + b) / gib, a / gib, na, b / gib, nb, fr / gib, tot / gib);
  return fail(XT_ENOMEM, buf);
}
}  // namespace

int xt_upload_peak_bytes(xt_ctx* c, double* bytes) {
  if(!c || !bytes) return fail(XT_EINVAL, "xt_upload_peak_bytes: null argument");
  *bytes = c->upload_peak;
  return XT_OK;
}

int xt_memory_required(xt_ctx* c, long nchol, double* bytes) {
  if(!c || !bytes) return fail(XT_EINVAL, "xt_memory_required: null argument");
  if(!c->have_plan) return fail(XT_ESTATE, "xt_memory_required: call xt_set_tiling first");
  const MemReq r = mem_req(c->plan.O, c->plan.V, nchol);
  bytes[0]       = r.operands;
  bytes[1]       = r.staging;
  bytes[2]       = r.split;
  return XT_OK;
}

// First non-finite element of each host input, checked before any device
// work (a NaN or an infinity would otherwise reach the energy sums).
static int check_finite(const char* fn, std::initializer_list<std::pair<const char*, std::pair<const double*, size_t>>> arrays) {
  for(const auto& a: arrays) {
    const double* p = a.second.first;
    const size_t  n = a.second.second;
    const long    chunks = (long) ((n + (1u << 22) - 1) >> 22);
    std::atomic<size_t> first{SIZE_MAX};
    parallel_for(chunks, [&](long k) {
      const size_t b = (size_t) k << 22, e = std::min(n, b + (1u << 22));
      for(size_t i = b; i < e; i++)
        if(!std::isfinite(p[i])) {
          size_t cur = first.load();
          while(i < cur && !first.compare_exchange_weak(cur, i)) {
          }
          break;
        }
    });
    if(first.load() != SIZE_MAX)
      return fail(XT_ERANGE, std::string(fn) + ": " + a.first + "[" + std::to_string(first.load()) +
                               "] is not finite");
  }
  return XT_OK;
}

}  // extern "C"

// Counted device allocations of an upload (track_malloc / track_free): the
// bytes live between the upload's own allocations and frees, and their
// maximum.  Allocations that rocBLAS makes for itself are not seen.
namespace {
struct UploadCount {
  std::unordered_map<void*, size_t> live;
  double                            bytes = 0.0, peak = 0.0;
};
thread_local UploadCount* g_count = nullptr;
// Opens a count on this thread for the lifetime of the object.
struct CountScope {
  UploadCount c;
  CountScope() { g_count = &c; }
  ~CountScope() { g_count = nullptr; }
};
}  // namespace

int xt::track_malloc(void** p, size_t bytes) {
  const hipError_t e = hipMalloc(p, bytes);
  if(e == hipSuccess && g_count) {
    g_count->live[*p] = bytes;
    g_count->bytes += (double) bytes;
    g_count->peak = std::max(g_count->peak, g_count->bytes);
  }
  return (int) e;
}

void xt::track_free(void* p) {
  if(!p) return;
  if(g_count) {
    auto it = g_count->live.find(p);
    if(it != g_count->live.end()) {
      g_count->bytes -= (double) it->second;
      g_count->live.erase(it);
    }
  }
  (void) hipFree(p);
}

extern "C" {

static int upload_common(xt_ctx* c, const double* eo, const double* ev, const double* t1,
                         const double* t2, const double* vvoo, const double* ovoo,
                         const double* vvvo) {
  const size_t o = c->plan.O, v = c->plan.V;
  XT_HIP(hipSetDevice(c->device));
  c->free_inputs();
  {
    const MemReq r  = mem_req(o, v, 0);
    const int    rc =
      check_free("xt_upload_closed_shell", r.operands, "operands", r.staging, "upload staging");
    if(rc) return rc;
  }
  hipStream_t st = c->stream;
  CountScope  count;
  c->upload_peak = 0.0;
  double *    t2_d = nullptr, *ovoo_d = nullptr, *vvvo_d = nullptr;
  auto        cleanup = [&]() {
    for(double* p: {t2_d, ovoo_d, vvvo_d})
      if(p) xt::track_free(p);
    c->free_inputs();
  };
  auto up = [&](double** dst, const double* src, size_t n) -> int {
    hipError_t e = (hipError_t) xt::track_malloc((void**) dst, n * sizeof(double));
    if(e != hipSuccess) {
      *dst = nullptr;
      return fail(XT_ENOMEM, "hipMalloc of " + std::to_string(n * sizeof(double)) + " bytes failed");
    }
    e = hipMemcpyAsync(*dst, src, n * sizeof(double), hipMemcpyHostToDevice, st);
    if(e != hipSuccess) return fail(XT_EHIP, std::string("upload copy: ") + hipGetErrorString(e));
    return XT_OK;
  };
  int rc;
  if((rc = up(&c->in.eo, eo, o)) || (rc = up(&c->in.ev, ev, v)) || (rc = up(&c->in.t1, t1, v * o)) ||
     (rc = up(&c->in.vvoo, vvoo, v * v * o * o)) || (rc = up(&t2_d, t2, v * v * o * o)) ||
     (rc = up(&ovoo_d, ovoo, o * v * o * o)) || (rc = up(&vvvo_d, vvvo, v * v * v * o))) {
    cleanup();
    return rc;
  }
  int err = xt::build_operands((int) o, (int) v, t2_d, ovoo_d, vvvo_d, c->in, st);
  if(err) {
    cleanup();
    return fail(err == (int) hipErrorOutOfMemory ? XT_ENOMEM : XT_EHIP,
                std::string("operand layout: ") + hipGetErrorString((hipError_t) err));
  }
  xt::track_free(t2_d);
  xt::track_free(ovoo_d);
  if(vvvo_d) xt::track_free(vvvo_d);
  c->upload_peak = count.c.peak;
  c->have_inputs = true;
  return XT_OK;
}

int xt_upload_cholesky(xt_ctx* c, const double* eps_occ, const double* eps_vir, const double* t1,
                       const double* t2, long nchol, const double* chol) {
  if(!c || !eps_occ || !eps_vir || !t1 || !t2 || !chol || nchol <= 0)
    return fail(XT_EINVAL, "xt_upload_cholesky: null argument or nchol <= 0");
  if(!c->have_plan) return fail(XT_ESTATE, "xt_upload_cholesky: call xt_set_tiling first");
  const size_t o = c->plan.O, v = c->plan.V, nmo = o + v;
  {
    const int rc = check_finite("xt_upload_cholesky",
                                {{"eps_occ", {eps_occ, o}}, {"eps_vir", {eps_vir, v}},
                                 {"t1", {t1, v * o}}, {"t2", {t2, v * v * o * o}},
                                 {"chol", {chol, nmo * nmo * (size_t) nchol}}});
    if(rc) return rc;
  }
  if(c->device < 0) return fail(XT_ENODEV, "xt_upload_cholesky: host-only context");
  if((long) (nmo * nchol) > (1l << 31) - 1 || (long) (v * o) > (1l << 31) - 1)
    return fail(XT_EINVAL, "xt_upload_cholesky: nmo * nchol beyond the BLAS 32-bit leading dimension");
  XT_HIP(hipSetDevice(c->device));
  c->free_inputs();
  {
    const MemReq r  = mem_req(o, v, nchol);
    const int    rc =
      check_free("xt_upload_cholesky", r.operands, "operands", r.staging, "upload staging");
    if(rc) return rc;
  }
  hipStream_t st = c->stream;
  CountScope  count;
  c->upload_peak = 0.0;
  double *    t2_d = nullptr, *ovoo_d = nullptr, *L_d = nullptr, *none = nullptr;
  auto        cleanup = [&]() {
    for(double* p: {t2_d, ovoo_d, L_d})
      if(p) xt::track_free(p);
    c->free_inputs();
  };
  auto up = [&](double** dst, const double* src, size_t n) -> int {
    hipError_t e = (hipError_t) xt::track_malloc((void**) dst, n * sizeof(double));
    if(e != hipSuccess) {
      *dst = nullptr;
      return fail(XT_ENOMEM, "hipMalloc of " + std::to_string(n * sizeof(double)) + " bytes failed");
    }
    if(src) e = hipMemcpyAsync(*dst, src, n * sizeof(double), hipMemcpyHostToDevice, st);
    if(e != hipSuccess) return fail(XT_EHIP, std::string("upload copy: ") + hipGetErrorString(e));
    return XT_OK;
  };
  const size_t vp = (v + 7) / 8 * 8;
  int          rc;
  if((rc = up(&c->in.eo, eps_occ, o)) || (rc = up(&c->in.ev, eps_vir, v)) ||
     (rc = up(&c->in.t1, t1, v * o)) || (rc = up(&t2_d, t2, v * v * o * o)) ||
     (rc = up(&L_d, chol, nmo * nmo * (size_t) nchol)) || (rc = up(&c->in.vvoo, nullptr, v * v * o * o)) ||
     (rc = up(&ovoo_d, nullptr, o * v * o * o)) || (rc = up(&c->in.xv, nullptr, v * v * o * vp))) {
    cleanup();
    return rc;
  }
  int err = xt::cholesky_integrals((int) o, (int) v, nchol, L_d, c->in.vvoo, ovoo_d, c->in.xv, st);
  if(!err) {
    xt::track_free(L_d);
    L_d = nullptr;
    err = xt::build_operands((int) o, (int) v, t2_d, ovoo_d, none, c->in, st);
  }
  if(!err) err = xt::build_xva((int) o, (int) v, c->in, st);
  if(err) {
    cleanup();
    return fail(err == (int) hipErrorOutOfMemory ? XT_ENOMEM : XT_EHIP,
                std::string("cholesky integrals: ") + hipGetErrorString((hipError_t) err));
  }
  xt::track_free(t2_d);
  xt::track_free(ovoo_d);
  c->upload_peak = count.c.peak;
  c->have_inputs = true;
  return XT_OK;
}

int xt_upload_closed_shell(xt_ctx* c, const double* eps_occ, const double* eps_vir,
                           const double* t1, const double* t2, const double* vvoo,
                           const double* ovoo, const double* vvvo) {
  if(!c || !eps_occ || !eps_vir || !t1 || !t2 || !vvoo || !ovoo || !vvvo)
    return fail(XT_EINVAL, "xt_upload_closed_shell: null argument");
  if(!c->have_plan) return fail(XT_ESTATE, "xt_upload_closed_shell: call xt_set_tiling first");
  {
    const size_t o = c->plan.O, v = c->plan.V;
    const int    rc = check_finite(
      "xt_upload_closed_shell",
      {{"eps_occ", {eps_occ, o}}, {"eps_vir", {eps_vir, v}}, {"t1", {t1, v * o}},
       {"t2", {t2, v * v * o * o}}, {"vvoo", {vvoo, v * v * o * o}},
       {"ovoo", {ovoo, o * v * o * o}}, {"vvvo", {vvvo, v * v * v * o}}});
    if(rc) return rc;
  }
  if(c->device < 0) return fail(XT_ENODEV, "xt_upload_closed_shell: host-only context");
  return upload_common(c, eps_occ, eps_vir, t1, t2, vvoo, ovoo, vvvo);
}

int xt_synth_upload(xt_ctx* c, uint64_t seed, double sigma, double t1_sigma) {
  if(!c) return fail(XT_EINVAL, "xt_synth_upload: null context");
  if(!c->have_plan) return fail(XT_ESTATE, "xt_synth_upload: call xt_set_tiling first");
  if(!std::isfinite(sigma) || !std::isfinite(t1_sigma))
    return fail(XT_ERANGE, "xt_synth_upload: sigma and t1_sigma must be finite");
  if(c->device < 0) return fail(XT_ENODEV, "xt_synth_upload: host-only context");
  const size_t o = c->plan.O, v = c->plan.V;
  std::vector<double> eo(o), ev(v), t1(v * o);
  std::vector<double> t2(v * v * o * o), vvoo(v * v * o * o), ovoo(o * v * o * o), vvvo(v * v * v * o);
  int rc = synth((int) o, (int) v, seed, sigma, t1_sigma, eo.data(), ev.data(), t1.data(),
                 t2.data(), vvoo.data(), ovoo.data(), vvvo.data());
  if(rc) return rc;
  return upload_common(c, eo.data(), ev.data(), t1.data(), t2.data(), vvoo.data(), ovoo.data(),
                       vvvo.data());
}

void xt_default_opts(xt_opts* o) {
  if(!o) return;
  std::memset(o, 0, sizeof(*o));
  o->precision        = XT_PREC_FP64;
  o->schedule         = XT_SCHED_DRIVER;
  o->use_graph        = 1;
  o->record_task_time = 0;
  o->workgroups       = 0;
}

// Device view of the mixed form's operands (the split must exist).
static xt::MixedDev mixed_dev(const xt_ctx* c) {
  const Plan&  P = c->plan;
  xt::MixedDev M = {};
  M.O            = P.O;
  M.V            = P.V;
  M.Ko           = (P.O + 15) / 16 * 16;
  M.Kv           = (P.V + 15) / 16 * 16;
  M.eo           = c->in.eo;
  M.ev           = c->in.ev;
  M.t1           = c->in.t1;
  M.vvoo         = c->in.vvoo;
  for(int j = 0; j < xt::kOperands; j++) {
    M.v[j]    = c->split.v[j];
    M.eblk[j] = c->split.eblk[j];
  }
  M.zero   = c->split.zero;
  M.pscale = c->split.pscale;
  M.nto    = P.noab / 2;
  M.ntv    = P.nvab / 2;
  return M;
}

static int ensure_plan_dev(xt_ctx* c) {
  const Plan& P  = c->plan;
  const long  nt = std::max<long>((long) P.dev.size(), 1);
  if(c->dev_tasks_n == (long) P.dev.size()) return XT_OK;
  c->free_plan_dev();
  XT_HIP(hipMalloc((void**) &c->d_tasks, sizeof(xt::TaskDev) * nt));
  XT_HIP(hipMalloc((void**) &c->d_qtask, sizeof(int) * nt));
  XT_HIP(hipMalloc((void**) &c->d_qtick, sizeof(long long) * nt));
  XT_HIP(hipMalloc((void**) &c->d_qbeg, sizeof(int) * (xt::kQueues + 1)));
  XT_HIP(hipMalloc((void**) &c->d_qtotal, sizeof(long long) * xt::kQueues));
  // the ticket counters and, after them, the run's error flag
  XT_HIP(hipMalloc((void**) &c->d_qctr, sizeof(unsigned long long) * (xt::kQueues + 1)));
  XT_HIP(hipMalloc((void**) &c->d_acc, sizeof(unsigned long long) * 4 * nt));
  XT_HIP(hipMalloc((void**) &c->d_cost, sizeof(unsigned long long) * nt));
  XT_HIP(hipMalloc((void**) &c->d_task_e, sizeof(double) * 2 * nt));
  XT_HIP(hipMalloc((void**) &c->d_e45, sizeof(double) * 2));
  if(!P.dev.empty())
    XT_HIP(hipMemcpy(c->d_tasks, P.dev.data(), sizeof(xt::TaskDev) * P.dev.size(),
                     hipMemcpyHostToDevice));
  c->dev_tasks_n = (long) P.dev.size();
  return XT_OK;
}

// Queues of one run.  Every task stays in one queue (its sub-tiles share
// operand panels in that XCD's L2).  Driver schedule: tasks dealt round-robin
// in driver order.  Cost schedules: longest-processing-time first onto the
// least-loaded queue, each queue processed in decreasing cost.
struct Queues {
  std::vector<int>       task, beg;
  std::vector<long long> tick, total;
};

// Tickets of task T under group shape glog (one group of 64 boxes per
// ceil(n/2^glog) cell of the task's box grid).
static long long task_tickets(const xt::TaskDev& T, const int* glog, bool mixed) {
  return xt::task_ticket_count(mixed ? T.nsubm : T.nsub, glog);
}

static Queues build_queues(const Plan& P, const std::vector<int>& sel, int schedule,
                           const std::vector<double>* cost, bool mixed, const int* glog) {
  const int                     nq = xt::kQueues;
  std::vector<std::vector<int>> per(nq);
  if(schedule == XT_SCHED_DRIVER) {
    for(size_t i = 0; i < sel.size(); i++) per[i % nq].push_back(sel[i]);
  }
  else {
    auto c = [&](int t) { return cost ? (*cost)[t] : P.dev[t].flops; };
    std::vector<int> order = sel;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return c(a) > c(b); });
    std::vector<double> load(nq, 0.0);
    for(int t: order) {
      const int q = (int) (std::min_element(load.begin(), load.end()) - load.begin());
      load[q] += c(t);
      per[q].push_back(t);
    }
  }
  Queues Q;
  Q.beg.push_back(0);
  for(int q = 0; q < nq; q++) {
    long long tk = 0;
    for(int t: per[q]) {
      Q.task.push_back(t);
      Q.tick.push_back(tk);
      tk += task_tickets(P.dev[t], glog, mixed);
    }
    Q.beg.push_back((int) Q.task.size());
    Q.total.push_back(tk);
  }
  return Q;
}

// While the run is in flight, prints the share of queue tickets the
// workgroups have taken (from the host-mapped mirror the kernels write; a
// copy on another stream would wait for the persistent kernel's CUs).
static int report_progress(xt_ctx* c, const xt_opts& o, const Queues& Q) {
  if(o.progress_ms <= 0) return XT_OK;
  long long total = 0;
  for(long long t: Q.total) total += t;
  const auto         t0   = std::chrono::steady_clock::now();
  auto               last = t0;
  for(;;) {
    // poll often so the caller regains control soon after the run drains
    std::this_thread::sleep_for(std::chrono::milliseconds(2));
    const hipError_t q = hipStreamQuery(c->stream);
    if(q == hipSuccess) break;
    if(q != hipErrorNotReady) return fail(XT_EHIP, std::string("progress: ") + hipGetErrorString(q));
    const auto now = std::chrono::steady_clock::now();
    if(now - last < std::chrono::milliseconds(o.progress_ms)) continue;
    last = now;
    long long taken = 0;
    for(int q2 = 0; q2 < xt::kQueues; q2++) {
      const unsigned long long v = __atomic_load_n(&c->h_prog[q2], __ATOMIC_RELAXED);
      taken += std::min<long long>((long long) v, Q.total[q2]);
    }
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::fprintf(stderr, "[xt] %.0f s: %lld of %lld tickets taken (%.1f%%)\n", s, taken, total,
                 total ? 100.0 * (double) taken / (double) total : 100.0);
    std::fflush(stderr);
  }
  return XT_OK;
}

// Launches the persistent kernel over the tasks `sel` (all if null), then the
// per-task fixed-point -> FP64 conversion and the fixed-order sum.
static int run_units(xt_ctx* c, const xt_opts& o, const std::vector<int>* tasks_sel) {
  const Plan&      P  = c->plan;
  const long       nt = (long) P.dev.size();
  std::vector<int> sel;
  if(tasks_sel) sel = *tasks_sel;
  else {
    sel.resize(nt);
    std::iota(sel.begin(), sel.end(), 0);
  }
  const std::vector<double>* learned =
    (o.schedule == XT_SCHED_LEARNED && (long) c->learned_cost.size() == nt) ? &c->learned_cost
                                                                            : nullptr;
  const bool  mixed = o.precision == XT_PREC_MIXED;
  const int*  glog  = mixed ? c->glog_mixed : c->glog_fp64;
  for(int t: sel)  // the kernels decode a task's tickets in 32-bit arithmetic
    if(task_tickets(P.dev[t], glog, mixed) >= (1ll << 31))
      return fail(XT_EINVAL, "run: a task of 2^31 or more box tickets");
  Queues      Q     = build_queues(P, sel, o.schedule, learned, mixed, glog);
  hipStream_t st = c->stream;
  if(!Q.task.empty()) {
    XT_HIP(hipMemcpyAsync(c->d_qtask, Q.task.data(), sizeof(int) * Q.task.size(),
                          hipMemcpyHostToDevice, st));
    XT_HIP(hipMemcpyAsync(c->d_qtick, Q.tick.data(), sizeof(long long) * Q.tick.size(),
                          hipMemcpyHostToDevice, st));
  }
  XT_HIP(hipMemcpyAsync(c->d_qbeg, Q.beg.data(), sizeof(int) * Q.beg.size(), hipMemcpyHostToDevice,
                        st));
  XT_HIP(hipMemcpyAsync(c->d_qtotal, Q.total.data(), sizeof(long long) * Q.total.size(),
                        hipMemcpyHostToDevice, st));
  xt::SpatialDev S;
  S.O    = P.O;
  S.V    = P.V;
  S.Op   = (P.O + 7) / 8 * 8;
  S.Vp   = (P.V + 7) / 8 * 8;
  S.eo   = c->in.eo;
  S.ev   = c->in.ev;
  S.t1   = c->in.t1;
  S.t2h  = c->in.t2h;
  S.t2ha = c->in.t2ha;
  S.ovk  = c->in.ovk;
  S.ovka = c->in.ovka;
  S.xv   = c->in.xv;
  S.xva  = c->in.xva;
  S.t2p  = c->in.t2p;
  S.t2pa = c->in.t2pa;
  S.vvoo = c->in.vvoo;
  xt::RunDev R;
  R.tasks   = c->d_tasks;
  R.qtask   = c->d_qtask;
  R.qtick   = c->d_qtick;
  R.qbeg    = c->d_qbeg;
  R.qtotal  = c->d_qtotal;
  R.qctr    = c->d_qctr;
  R.err     = c->d_qctr + xt::kQueues;
  R.acc     = c->d_acc;
  R.cost    = o.record_task_time ? c->d_cost : nullptr;
  R.prog    = c->d_prog;
  for(int q = 0; q < xt::kQueues; q++) __atomic_store_n(&c->h_prog[q], 0ull, __ATOMIC_RELAXED);
  R.n_tasks = (int) nt;
  for(int p = 0; p < 6; p++) R.glog[p] = glog[p];
  long long tickets = 0;
  for(long long t: Q.total) tickets += t;
  // the FP64 kernel's LDS and registers fit three workgroups on a CU, the
  // mixed kernel's one
  int grid = o.workgroups > 0 ? o.workgroups : c->n_cu * 3;
  grid     = (int) std::max<long long>(1, std::min<long long>(grid, tickets));
  xt::MixedDev M = {};
  if(mixed) {
    M    = mixed_dev(c);
    grid = o.workgroups > 0 ? o.workgroups : c->n_cu;
    grid = (int) std::max<long long>(1, std::min<long long>(grid, tickets));
  }
  // the launch sequence; counters and accumulators are cleared by fill
  // kernels (no memset nodes when captured)
  auto body = [&](bool events) -> int {
    int err = xt::launch_zero_u64(c->d_qctr, xt::kQueues + 1, st);
    if(!err) err = xt::launch_zero_u64(c->d_acc, 4 * std::max<long>(nt, 1), st);
    if(!err) err = xt::launch_zero_u64(c->d_cost, std::max<long>(nt, 1), st);
    if(err) return fail(XT_EHIP, std::string("zero: ") + hipGetErrorString((hipError_t) err));
    if(events) XT_HIP(hipEventRecord(c->ev0, st));
    err = mixed ? xt::launch_spatial_mixed(M, R, grid, st) : xt::launch_spatial_fp64(S, R, grid, st);
    if(err) return fail(XT_EHIP, std::string("launch: ") + hipGetErrorString((hipError_t) err));
    if(events) XT_HIP(hipEventRecord(c->ev1, st));
    err = xt::launch_reduce(c->d_tasks, (int) nt, c->d_acc, c->d_task_e, c->d_e45, st);
    if(err) return fail(XT_EHIP, std::string("reduce: ") + hipGetErrorString((hipError_t) err));
    return XT_OK;
  };
  if(!o.use_graph) {
    int rc = body(true);
    return rc ? rc : report_progress(c, o, Q);
  }
  xt_ctx::GraphKey key;
  key.precision = o.precision;
  key.grid      = grid;
  key.timed     = o.record_task_time ? 1 : 0;
  key.n_tasks   = nt;
  key.inputs    = mixed ? (const void*) c->split.v[0] : (const void*) c->in.xv;
  if(!c->graph || !(c->graph_key == key)) {
    c->drop_graph();
    XT_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int        rc = body(false);
    hipGraph_t g  = nullptr;
    hipError_t e  = hipStreamEndCapture(st, &g);
    if(rc) {
      if(g) (void) hipGraphDestroy(g);
      return rc;
    }
    if(e != hipSuccess) return fail(XT_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
    (void) hipGraphDestroy(g);
    if(e != hipSuccess) {
      c->graph = nullptr;
      return fail(XT_EHIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    c->graph_key = key;
  }
  // with a graph the timed span also holds the two small fill and reduce kernels
  XT_HIP(hipEventRecord(c->ev0, st));
  XT_HIP(hipGraphLaunch(c->graph, st));
  XT_HIP(hipEventRecord(c->ev1, st));
  return report_progress(c, o, Q);
}

// Summed workgroup time of each task (ms), from wall-clock ticks.
static int task_times(xt_ctx* c, long nt, float* task_ms) {
  std::vector<unsigned long long> ticks(nt);
  XT_HIP(hipMemcpy(ticks.data(), c->d_cost, nt * 8, hipMemcpyDeviceToHost));
  int rate_khz = 100000;
  (void) hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, c->device);
  for(long k = 0; k < nt; k++) task_ms[k] = (float) ((double) ticks[k] / rate_khz);
  return XT_OK;
}

// Error flag of the last run (after its stream has drained): a box whose
// energy partial was not finite or beyond 2^62 Eh did not enter the exact
// fixed-point sums, so the run's energies are not returned.
static int check_run_flag(xt_ctx* c, const char* fn) {
  unsigned long long flag = 0;
  XT_HIP(hipMemcpy(&flag, c->d_qctr + xt::kQueues, sizeof(flag), hipMemcpyDeviceToHost));
  if(flag)
    return fail(XT_ERANGE, std::string(fn) +
                             ": an energy partial was not finite or beyond 2^62 Eh (zero or "
                             "non-finite orbital-energy denominators, or non-finite inputs)");
  return XT_OK;
}

static int check_run_state(xt_ctx* c, const xt_opts& o) {
  if(c->device < 0) return fail(XT_ENODEV, "xt_run: host-only context");
  if(!c->have_plan || !c->have_inputs) return fail(XT_ESTATE, "xt_run: tiling or inputs missing");
  if(o.precision != XT_PREC_FP64 && o.precision != XT_PREC_MIXED)
    return fail(XT_EINVAL, "xt_run: unknown precision");
  if(o.precision == XT_PREC_MIXED && !c->have_split) {
    XT_HIP(hipSetDevice(c->device));
    // spatial tile of every orbital (the beta half of the tiling mirrors alpha)
    const Plan&      P = c->plan;
    std::vector<int> to(P.O), tv(P.V);
    for(int t = 0; t < P.noab / 2; t++)
      for(int i = 0; i < P.k_range[t]; i++) to[P.spatial_off[t] + i] = t;
    for(int t = 0; t < P.nvab / 2; t++)
      for(int i = 0; i < P.k_range[P.noab + t]; i++) tv[P.spatial_off[P.noab + t] + i] = t;
    {
      const MemReq r  = mem_req(P.O, P.V, 0);
      const int    rc =
        check_free("xt_run (mixed precision)", r.split, "FP16 split copies", 0.0, "other");
      if(rc) return rc;
    }
    int err = xt::build_split(c->in, P.O, P.V, to.data(), P.noab / 2, tv.data(), P.nvab / 2,
                              c->split, c->stream);
    if(err) {
      c->free_split();
      return fail(err == (int) hipErrorOutOfMemory ? XT_ENOMEM : XT_EHIP,
                  std::string("split operands: ") + hipGetErrorString((hipError_t) err));
    }
    c->have_split = true;
  }
  if(o.schedule < XT_SCHED_DRIVER || o.schedule > XT_SCHED_LEARNED)
    return fail(XT_EINVAL, "xt_run: bad schedule");
  XT_HIP(hipSetDevice(c->device));
  int rc = ensure_plan_dev(c);
  if(rc) return rc;
  const long nt = (long) c->plan.dev.size();
  if(o.precision == XT_PREC_MIXED && c->split.pscale_n != nt) {
    if(c->split.pscale) (void) hipFree(c->split.pscale);
    c->split.pscale   = nullptr;
    c->split.pscale_n = 0;
    XT_HIP(hipMalloc((void**) &c->split.pscale, sizeof(float) * 18 * std::max<long>(nt, 1)));
    const xt::MixedDev M = mixed_dev(c);
    int err = xt::build_part_scales(c->d_tasks, (int) nt, M, c->split.pscale, c->stream);
    if(!err) err = (int) hipStreamSynchronize(c->stream);
    if(err) return fail(XT_EHIP, std::string("part scales: ") + hipGetErrorString((hipError_t) err));
    c->split.pscale_n = nt;
  }
  return XT_OK;
}

int xt_prepare(xt_ctx* c, int precision) {
  if(!c) return fail(XT_EINVAL, "xt_prepare: null context");
  xt_opts o;
  xt_default_opts(&o);
  o.precision = precision;
  return check_run_state(c, o);
}

int xt_run(xt_ctx* c, const xt_opts* opts, double* e45, double* task_partials, float* task_ms) {
  if(!c || !e45) return fail(XT_EINVAL, "xt_run: null argument");
  xt_opts o;
  if(opts) o = *opts;
  else xt_default_opts(&o);
  int rc = check_run_state(c, o);
  if(rc) return rc;
  const long nt = (long) c->plan.dev.size();
  if((rc = run_units(c, o, nullptr))) return rc;
  hipStream_t st = c->stream;
  XT_HIP(hipMemcpyAsync(e45, c->d_e45, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  if(task_partials)
    XT_HIP(hipMemcpyAsync(task_partials, c->d_task_e, 2 * nt * sizeof(double),
                          hipMemcpyDeviceToHost, st));
  XT_HIP(hipStreamSynchronize(st));
  float ms = 0;
  XT_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_ms = ms;
  if((rc = check_run_flag(c, "xt_run"))) {
    e45[0] = e45[1] = std::nan("");
    if(task_partials) std::fill(task_partials, task_partials + 2 * nt, std::nan(""));
    return rc;
  }
  if(o.record_task_time && task_ms) return task_times(c, nt, task_ms);
  return XT_OK;
}

int xt_sum_task_energies(long n, const double* task_partials, double* e45) {
  if(n < 0 || !e45 || (n > 0 && !task_partials))
    return fail(XT_EINVAL, "xt_sum_task_energies: null argument or negative count");
  if(n > INT_MAX) return fail(XT_EINVAL, "xt_sum_task_energies: more than INT_MAX tasks");
  // the order of k_reduce_final: kReduceThreads runs of ceil(n / kReduceThreads)
  // consecutive tasks, each summed in task order, then a halving tree
  constexpr int kT = xt::kReduceThreads;
  double        ra[kT], rb[kT];
  const long    per = (n + kT - 1) / kT;
  for(int i = 0; i < kT; i++) {
    double     a = 0.0, b = 0.0;
    const long lo = i * per, hi = std::min(n, lo + per);
    for(long t = lo; t < hi; t++) {
      a += task_partials[2 * t];
      b += task_partials[2 * t + 1];
    }
    ra[i] = a;
    rb[i] = b;
  }
  for(int s = kT / 2; s >= 1; s >>= 1)
    for(int i = 0; i < s; i++) {
      ra[i] += ra[i + s];
      rb[i] += rb[i + s];
    }
  e45[0] = ra[0];
  e45[1] = rb[0];
  return XT_OK;
}

int xt_run_tasks(xt_ctx* c, const xt_opts* opts, long n, const long* which, double* e45,
                 double* task_partials, float* task_ms) {
  if(!c || !which || n <= 0 || !e45) return fail(XT_EINVAL, "xt_run_tasks: bad argument");
  xt_opts o;
  if(opts) o = *opts;
  else xt_default_opts(&o);
  int rc = check_run_state(c, o);
  if(rc) return rc;
  const long       nt = (long) c->plan.dev.size();
  std::vector<int> sel(n);
  std::vector<char> seen(nt, 0);
  for(long i = 0; i < n; i++) {
    if(which[i] < 0 || which[i] >= nt) return fail(XT_EINVAL, "xt_run_tasks: task out of range");
    if(seen[which[i]]) return fail(XT_EINVAL, "xt_run_tasks: repeated task");
    seen[which[i]] = 1;
    sel[i]         = (int) which[i];
  }
  if((rc = run_units(c, o, &sel))) return rc;
  std::vector<double> te(2 * nt);
  XT_HIP(hipMemcpyAsync(te.data(), c->d_task_e, 2 * nt * sizeof(double), hipMemcpyDeviceToHost,
                        c->stream));
  XT_HIP(hipStreamSynchronize(c->stream));
  float ms = 0;
  XT_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_ms = ms;
  if((rc = check_run_flag(c, "xt_run_tasks"))) {
    e45[0] = e45[1] = std::nan("");
    if(task_partials) std::fill(task_partials, task_partials + 2 * n, std::nan(""));
    return rc;
  }
  e45[0] = e45[1] = 0.0;
  for(long i = 0; i < n; i++) {  // listed order
    e45[0] += te[2 * which[i]];
    e45[1] += te[2 * which[i] + 1];
    if(task_partials) {
      task_partials[2 * i]     = te[2 * which[i]];
      task_partials[2 * i + 1] = te[2 * which[i] + 1];
    }
  }
  if(o.record_task_time && task_ms) {
    std::vector<float> all(nt);
    if((rc = task_times(c, nt, all.data()))) return rc;
    for(long i = 0; i < n; i++) task_ms[i] = all[which[i]];
  }
  return XT_OK;
}

int xt_queue_layout(xt_ctx* c, const xt_opts* opts, long n, const long* which, int* qbeg,
                    int* qtask, long long* qtick, long long* qtotal, int* nbox, int* glog) {
  if(!c || !opts || n < 0 || (n > 0 && !which) || !qbeg || !qtask || !qtick || !qtotal || !nbox ||
     !glog)
    return fail(XT_EINVAL, "xt_queue_layout: bad argument");
  if(!c->have_plan) return fail(XT_ESTATE, "xt_queue_layout: no tiling");
  if(opts->precision != XT_PREC_FP64 && opts->precision != XT_PREC_MIXED)
    return fail(XT_EINVAL, "xt_queue_layout: unknown precision");
  const long       nt = (long) c->plan.dev.size();
  std::vector<int> sel;
  if(n == 0) {
    sel.resize(nt);
    std::iota(sel.begin(), sel.end(), 0);
  }
  else {
    std::vector<char> seen(nt, 0);
    for(long i = 0; i < n; i++) {
      if(which[i] < 0 || which[i] >= nt) return fail(XT_EINVAL, "xt_queue_layout: task out of range");
      if(seen[which[i]]) return fail(XT_EINVAL, "xt_queue_layout: repeated task");
      seen[which[i]] = 1;
      sel.push_back((int) which[i]);
    }
  }
  // the same choice of costs and group shape as run_units
  const std::vector<double>* learned =
    (opts->schedule == XT_SCHED_LEARNED && (long) c->learned_cost.size() == nt) ? &c->learned_cost
                                                                               : nullptr;
  const bool mixed = opts->precision == XT_PREC_MIXED;
  const int* g     = mixed ? c->glog_mixed : c->glog_fp64;
  Queues     Q     = build_queues(c->plan, sel, opts->schedule, learned, mixed, g);
  std::copy(Q.beg.begin(), Q.beg.end(), qbeg);
  std::copy(Q.task.begin(), Q.task.end(), qtask);
  std::copy(Q.tick.begin(), Q.tick.end(), qtick);
  std::copy(Q.total.begin(), Q.total.end(), qtotal);
  for(long t = 0; t < nt; t++)
    for(int p = 0; p < 6; p++)
      nbox[6 * t + p] = mixed ? c->plan.dev[t].nsubm[p] : c->plan.dev[t].nsub[p];
  std::copy(g, g + 6, glog);
  return XT_OK;
}

int xt_task_codes(xt_ctx* c, long task, uint64_t* codes) {
  if(!c || !codes) return fail(XT_EINVAL, "xt_task_codes: null argument");
  if(!c->have_plan) return fail(XT_ESTATE, "xt_task_codes: no tiling");
  if(task < 0 || task >= (long) c->plan.dev.size())
    return fail(XT_EINVAL, "xt_task_codes: task out of range");
  codes[0] = c->plan.dev[task].d1code;
  codes[1] = c->plan.dev[task].d2code;
  codes[2] = c->plan.dev[task].s1code;
  return XT_OK;
}

// ----------------------------------------------------- device-side gather
static xt::GatherTiling gather_tiling(const xt_ctx* c) {
  const Plan& P = c->plan;
  return xt::GatherTiling{P.noab, P.nvab, P.k_range.data(), P.k_spin.data(), P.spatial_off.data()};
}

int xt_packed_sizes(xt_ctx* c, long* sizes) {
  if(!c || !sizes) return fail(XT_EINVAL, "xt_packed_sizes: null argument");
  if(!c->have_plan) return fail(XT_ESTATE, "xt_packed_sizes: no tiling");
  xt::packed_sizes(gather_tiling(c), sizes);
  return XT_OK;
}

int xt_gather_task(xt_ctx* c, const long* task, void* stream, double* d1_t2, double* d1_v2,
                   double* d2_t2, double* d2_v2, double* s1_t1, double* s1_v2, int* s1_size,
                   int* s1_exec, int* d1_size, int* d1_exec, int* d2_size, int* d2_exec,
                   int* counts) {
  if(!c || !task || !d1_t2 || !d1_v2 || !d2_t2 || !d2_v2 || !s1_t1 || !s1_v2 || !s1_size ||
     !s1_exec || !d1_size || !d1_exec || !d2_size || !d2_exec || !counts)
    return fail(XT_EINVAL, "xt_gather_task: null argument");
  if(!c->have_plan) return fail(XT_ESTATE, "xt_gather_task: no tiling");
  if(c->device < 0) return fail(XT_ENODEV, "xt_gather_task: host-only context");
  if(!c->have_inputs) return fail(XT_ESTATE, "xt_gather_task: no inputs uploaded");
  const Plan& P = c->plan;
  for(int k = 0; k < 6; k++) {
    const bool occ = k < 3;
    const long lo = occ ? 0 : P.noab, hi = occ ? P.noab : P.noab + P.nvab;
    if(task[k] < lo || task[k] >= hi) return fail(XT_EINVAL, "xt_gather_task: tile index out of range");
  }
  xt::plan_gather(gather_tiling(c), P.restricted, task, c->h_gather, s1_size, s1_exec, d1_size,
                  d1_exec, d2_size, d2_exec, counts);
  const long n = (long) c->h_gather.size();
  c->gather_elems = 0;
  for(const auto& B: c->h_gather) {
    if(B.n >= (1l << 31)) return fail(XT_EINVAL, "xt_gather_task: block of 2^31 or more doubles");
    c->gather_elems += B.n;
  }
  if(n == 0) return XT_OK;
  XT_HIP(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t) stream;
  if(n > c->d_gather_cap) {
    // the previous descriptors may still be read by a gather on the stream
    XT_HIP(hipStreamSynchronize(st));
    if(c->d_gather) (void) hipFree(c->d_gather);
    c->d_gather     = nullptr;
    c->d_gather_cap = 0;
    XT_HIP(hipMalloc((void**) &c->d_gather, sizeof(xt::GatherBlock) * n));
    c->d_gather_cap = n;
  }
  XT_HIP(hipMemcpyAsync(c->d_gather, c->h_gather.data(), sizeof(xt::GatherBlock) * n,
                        hipMemcpyHostToDevice, st));
  long max_elems = 0;
  for(const auto& B: c->h_gather) max_elems = std::max(max_elems, B.n);
  xt::GatherSrc S;
  S.O    = P.O;
  S.V    = P.V;
  S.Op   = (P.O + 7) / 8 * 8;
  S.Vp   = (P.V + 7) / 8 * 8;
  S.t1   = c->in.t1;
  S.t2h  = c->in.t2h;
  S.vvoo = c->in.vvoo;
  S.ovk  = c->in.ovk;
  S.xv   = c->in.xv;
  xt::GatherDst D = {{d1_t2, d1_v2, d2_t2, d2_v2, s1_t1, s1_v2}};
  const int     err = xt::launch_gather(c->d_gather, (int) n, max_elems, S, D, st);
  if(err) return fail(XT_EHIP, std::string("xt_gather_task: ") + hipGetErrorString((hipError_t) err));
  return XT_OK;
}

int xt_expand_amplitudes(xt_ctx* c, void* stream, double* t1_so, double* t2_so) {
  if(!c) return fail(XT_EINVAL, "xt_expand_amplitudes: null context");
  if(!c->have_plan) return fail(XT_ESTATE, "xt_expand_amplitudes: no tiling");
  if(c->device < 0) return fail(XT_ENODEV, "xt_expand_amplitudes: host-only context");
  if(!c->have_inputs) return fail(XT_ESTATE, "xt_expand_amplitudes: no inputs uploaded");
  if(!t1_so && !t2_so) return XT_OK;
  XT_HIP(hipSetDevice(c->device));
  const Plan&   P = c->plan;
  xt::GatherSrc S = {};
  S.O             = P.O;
  S.V             = P.V;
  S.Op            = (P.O + 7) / 8 * 8;
  S.Vp            = (P.V + 7) / 8 * 8;
  S.t1            = c->in.t1;
  S.t2h           = c->in.t2h;
  const int err   = xt::launch_expand_t(S, t1_so, t2_so, stream);
  if(err) return fail(XT_EHIP, std::string("xt_expand_amplitudes: ") + hipGetErrorString((hipError_t) err));
  return XT_OK;
}

int xt_gather_last_elems(xt_ctx* c, long* n) {
  if(!c || !n) return fail(XT_EINVAL, "xt_gather_last_elems: null argument");
  *n = c->gather_elems;
  return XT_OK;
}

int xt_last_kernel_ms(xt_ctx* c, double* ms) {
  if(!c || !ms) return fail(XT_EINVAL, "xt_last_kernel_ms: null argument");
  *ms = c->last_ms;
  return XT_OK;
}

// ----------------------------------------------------------- autotuner
namespace {
const int kTuneShapes[XT_TUNE_CANDIDATES][6] = {
  {1, 1, 1, 1, 1, 1},  // two boxes along every position
  {0, 0, 0, 2, 2, 2},  // 4x4x4 over the particles
  {2, 2, 2, 0, 0, 0},  // 4x4x4 over the holes
  {0, 1, 1, 1, 1, 2},
  {1, 1, 1, 0, 1, 2},
  {0, 0, 2, 2, 2, 0},
};

bool valid_shape(const int* g) {
  int sum = 0;
  for(int p = 0; p < 6; p++) {
    if(g[p] < 0 || g[p] > xt::kGroupLog2) return false;
    sum += g[p];
  }
  return sum == xt::kGroupLog2;
}
}  // namespace

int xt_set_group_shape(xt_ctx* c, int precision, const int* glog) {
  if(!c || !glog) return fail(XT_EINVAL, "xt_set_group_shape: null argument");
  if(precision != XT_PREC_FP64 && precision != XT_PREC_MIXED)
    return fail(XT_EINVAL, "xt_set_group_shape: unknown precision");
  if(!valid_shape(glog)) return fail(XT_EINVAL, "xt_set_group_shape: extents must sum to 2^6");
  int* dst = precision == XT_PREC_MIXED ? c->glog_mixed : c->glog_fp64;
  std::copy(glog, glog + 6, dst);
  c->drop_graph();  // the captured launch holds the shape
  return XT_OK;
}

int xt_get_group_shape(xt_ctx* c, int precision, int* glog) {
  if(!c || !glog) return fail(XT_EINVAL, "xt_get_group_shape: null argument");
  if(precision != XT_PREC_FP64 && precision != XT_PREC_MIXED)
    return fail(XT_EINVAL, "xt_get_group_shape: unknown precision");
  const int* src = precision == XT_PREC_MIXED ? c->glog_mixed : c->glog_fp64;
  std::copy(src, src + 6, glog);
  return XT_OK;
}

int xt_autotune(xt_ctx* c, const xt_opts* opts, long n_sample, int* best, double* times_ms,
                int* shapes) {
  if(!c || !opts || n_sample <= 0) return fail(XT_EINVAL, "xt_autotune: bad argument");
  if(!c->have_plan) return fail(XT_ESTATE, "xt_autotune: no tiling");
  const long n = (long) c->plan.dev.size();
  if(n == 0) return fail(XT_ESTATE, "xt_autotune: no tasks");
  // tasks spread evenly over the flop-sorted list
  std::vector<long> order(n);
  std::iota(order.begin(), order.end(), 0L);
  std::stable_sort(order.begin(), order.end(),
                   [&](long a, long b) { return c->plan.dev[a].flops < c->plan.dev[b].flops; });
  const long        ns = std::min(n_sample, n);
  std::vector<long> pick(ns);
  for(long i = 0; i < ns; i++) pick[i] = order[ns == 1 ? n / 2 : (i * (n - 1)) / (ns - 1)];
  std::sort(pick.begin(), pick.end());
  pick.erase(std::unique(pick.begin(), pick.end()), pick.end());
  xt_opts o          = *opts;
  o.record_task_time = 0;
  o.progress_ms      = 0;
  const int prec     = o.precision;
  int       rc       = XT_OK;
  std::vector<double> e45(2), part(2 * pick.size());
  double              best_ms = 0;
  int                 best_i  = -1;
  for(int k = -1; k < XT_TUNE_CANDIDATES; k++) {  // k = -1: warm-up run
    if((rc = xt_set_group_shape(c, prec, kTuneShapes[k < 0 ? 0 : k]))) return rc;
    if((rc = xt_run_tasks(c, &o, (long) pick.size(), pick.data(), e45.data(), part.data(), nullptr)))
      return rc;
    if(k < 0) continue;
    const double ms = c->last_ms;
    if(times_ms) times_ms[k] = ms;
    if(shapes) std::copy(kTuneShapes[k], kTuneShapes[k] + 6, shapes + 6 * k);
    if(best_i < 0 || ms < best_ms) {
      best_ms = ms;
      best_i  = k;
    }
  }
  if(best) *best = best_i;
  return xt_set_group_shape(c, prec, kTuneShapes[best_i]);
}

// ----------------------------------------------------- learned schedule
int xt_task_features(xt_ctx* c, int precision, float* feats) {
  if(!c || !feats) return fail(XT_EINVAL, "xt_task_features: null argument");
  if(!c->have_plan) return fail(XT_ESTATE, "xt_task_features: no tiling");
  if(precision != XT_PREC_FP64 && precision != XT_PREC_MIXED)
    return fail(XT_EINVAL, "xt_task_features: unknown precision");
  for(size_t k = 0; k < c->plan.dev.size(); k++) {
    for(int p = 0; p < 6; p++) feats[7 * k + p] = (float) c->plan.dev[k].dim[p] / 48.f;
    feats[7 * k + 6] = precision == XT_PREC_MIXED ? 1.f : 0.f;
  }
  return XT_OK;
}

int xt_set_task_cost(xt_ctx* c, const double* cost) {
  if(!c || !cost) return fail(XT_EINVAL, "xt_set_task_cost: null argument");
  if(!c->have_plan) return fail(XT_ESTATE, "xt_set_task_cost: no tiling");
  const size_t n = c->plan.dev.size();
  for(size_t k = 0; k < n; k++)
    if(!(cost[k] > 0.0) || !std::isfinite(cost[k]))
      return fail(XT_EINVAL, "xt_set_task_cost: costs must be positive and finite");
  c->learned_cost.assign(cost, cost + n);
  return XT_OK;
}

int xt_learn_schedule(xt_ctx* c, xt_mlp* m, const xt_opts* opts, long n_sample, int precision_mask,
                      int steps, float lr, xt_learn_stats* stats) {
  if(!c || !m || !opts || n_sample <= 0 || !(precision_mask & 3) || steps <= 0 || !(lr > 0.f))
    return fail(XT_EINVAL, "xt_learn_schedule: bad argument");
  if(!c->have_plan) return fail(XT_ESTATE, "xt_learn_schedule: no tiling");
  const long n = (long) c->plan.dev.size();
  if(n == 0) return fail(XT_ESTATE, "xt_learn_schedule: no tasks");
  const auto t0 = std::chrono::steady_clock::now();
  // tasks spread evenly over the flop-sorted list
  std::vector<long> order(n);
  std::iota(order.begin(), order.end(), 0L);
  std::stable_sort(order.begin(), order.end(),
                   [&](long a, long b) { return c->plan.dev[a].flops < c->plan.dev[b].flops; });
  const long        ns = std::min(n_sample, n);
  std::vector<long> pick(ns);
  for(long i = 0; i < ns; i++) pick[i] = order[ns == 1 ? n / 2 : (i * (n - 1)) / (ns - 1)];
  std::sort(pick.begin(), pick.end());
  pick.erase(std::unique(pick.begin(), pick.end()), pick.end());
  // timing runs
  std::vector<float>  x, t;
  std::vector<double> lt, lf;  // log time, log flops of each sample
  std::vector<float>  feats(7 * n);
  for(int prec = 0; prec < 2; prec++) {
    if(!(precision_mask & (1 << prec))) continue;
    xt_opts o          = *opts;
    o.precision        = prec == 0 ? XT_PREC_FP64 : XT_PREC_MIXED;
    o.record_task_time = 1;
    o.progress_ms      = 0;
    std::vector<double> e45(2), part(2 * pick.size());
    std::vector<float>  ms(pick.size());
    int rc = xt_run_tasks(c, &o, (long) pick.size(), pick.data(), e45.data(), part.data(), ms.data());
    if(rc) return rc;
    if((rc = xt_task_features(c, o.precision, feats.data()))) return rc;
    for(size_t i = 0; i < pick.size(); i++) {
      if(!(ms[i] > 0.f)) continue;
      const double fl = c->plan.dev[pick[i]].flops;
      for(int j = 0; j < 7; j++) x.push_back(feats[7 * pick[i] + j]);
      lt.push_back(std::log((double) ms[i]));
      lf.push_back(std::log(fl));
    }
  }
  const long   nt = (long) lt.size();
  if(nt < 2) return fail(XT_ESTATE, "xt_learn_schedule: too few timed tasks");
  const auto   t1 = std::chrono::steady_clock::now();
  // standardised target: log time per flop
  double mu = 0, sd = 0;
  for(long i = 0; i < nt; i++) mu += lt[i] - lf[i];
  mu /= nt;
  for(long i = 0; i < nt; i++) sd += (lt[i] - lf[i] - mu) * (lt[i] - lf[i] - mu);
  sd = std::sqrt(sd / nt);
  if(!(sd > 1e-12)) sd = 1.0;
  t.resize(nt);
  for(long i = 0; i < nt; i++) t[i] = (float) ((lt[i] - lf[i] - mu) / sd);
  std::vector<float> hist(steps);
  int rc = xt_mlp_train(m, nt, x.data(), t.data(), steps, lr, hist.data());
  if(rc) return rc;
  std::vector<float> fit(nt);
  if((rc = xt_mlp_predict(m, nt, x.data(), fit.data()))) return rc;
  double se_m = 0, se_f = 0;
  for(long i = 0; i < nt; i++) {
    const double pm = lf[i] + mu + sd * fit[i], pf = lf[i] + mu;
    se_m += (pm - lt[i]) * (pm - lt[i]);
    se_f += (pf - lt[i]) * (pf - lt[i]);
  }
  // costs of every task at the scheduled precision
  if((rc = xt_task_features(c, opts->precision, feats.data()))) return rc;
  std::vector<float> y(n);
  if((rc = xt_mlp_predict(m, n, feats.data(), y.data()))) return rc;
  std::vector<double> cost(n);
  for(long k = 0; k < n; k++) {
    const double fl = c->plan.dev[k].flops;
    const double cc = fl * std::exp(mu + sd * (double) y[k]);
    cost[k]         = (std::isfinite(cc) && cc > 0.0) ? cc : fl;
  }
  c->learned_cost = cost;
  const auto t2   = std::chrono::steady_clock::now();
  if(stats) {
    stats->n_sample       = (long) pick.size();
    stats->final_loss     = hist.back();
    stats->fit_rmse_model = std::sqrt(se_m / nt);
    stats->fit_rmse_flops = std::sqrt(se_f / nt);
    stats->calib_ms       = std::chrono::duration<double, std::milli>(t1 - t0).count();
    stats->train_ms       = std::chrono::duration<double, std::milli>(t2 - t1).count();
  }
  return XT_OK;
}

// ------------------------------------------------------ per-task drop-in
namespace {
using xt::kMaxNoab;
using xt::kMaxNvab;
constexpr int kMetaInts = xt::kPackedMetaInts;
constexpr int kRing     = 4;

struct PackedState {
  int        device = -1;
  int*       pinned = nullptr;
  int*       dmeta  = nullptr;
  hipEvent_t ev[kRing] = {};
  int        slot      = 0;
  ~PackedState() {
    // Resources live until thread exit; the runtime may already be torn down
    // at process exit, so errors are ignored.
    if(device >= 0) {
      for(auto& e: ev)
        if(e) (void) hipEventDestroy(e);
      if(pinned) (void) hipHostFree(pinned);
      if(dmeta) (void) hipFree(dmeta);
    }
  }
};
thread_local PackedState g_packed;

// Arguments of fully_fused_ccsd_t_gpu<double> (nontc:2800-2817) as the two
// entry points receive them.
struct PackedCall {
  size_t        num_blocks, d[6];
  const double *d1_t2, *d1_v2, *d2_t2, *d2_v2, *s1_t1, *s1_v2;
  const int *   d1_size, *d1_exec, *d2_size, *d2_exec, *s1_size, *s1_exec;
  size_t        noab, max_d1_t2, max_d1_v2, nvab, max_d2_t2, max_d2_v2, max_s1_t1, max_s1_v2;
  const double* evl[6];
  double*       partial;
};

// Checks a launcher call, packs its host size/exec arrays into meta
// (kMetaInts ints: d1_size, d1_exec, d2_size, d2_exec, s1_size, s1_exec) and
// fills the kernel arguments except the device copy of meta.
int packed_prepare(const char* fn, const PackedCall& P, int* meta, int* n_meta, xt::PackedArgs& A) {
  const std::string f(fn);
  if(!P.d1_t2 || !P.d1_v2 || !P.d2_t2 || !P.d2_v2 || !P.s1_t1 || !P.s1_v2 || !P.d1_size ||
     !P.d1_exec || !P.d2_size || !P.d2_exec || !P.s1_size || !P.s1_exec || !P.evl[0] ||
     !P.evl[1] || !P.evl[2] || !P.evl[3] || !P.evl[4] || !P.evl[5] || !P.partial)
    return fail(XT_EINVAL, f + ": null argument");
  if(P.noab == 0 || P.nvab == 0 || P.noab > (size_t) kMaxNoab || P.nvab > (size_t) kMaxNvab)
    return fail(XT_EINVAL, f + ": noab/nvab outside 1..50 / 1..140");
  size_t nb = 1;
  for(size_t x: P.d) {
    if(x == 0 || x > 4096) return fail(XT_EINVAL, f + ": bad tile size");
    nb *= (x + 3) / 4;
  }
  if(nb != P.num_blocks) return fail(XT_EINVAL, f + ": num_blocks != prod(ceil(dim/4))");
  for(int i = 0; i < 6; i++)
    if((size_t) P.s1_size[i] != P.d[i]) return fail(XT_EINVAL, f + ": s1_size disagrees with tile sizes");
  for(size_t b = 0; b < P.noab; b++)
    if(P.d1_size[7 * b + 3] <= 0 || P.d1_size[7 * b + 3] > 4096)
      return fail(XT_EINVAL, f + ": bad h7 tile size");
  for(size_t b = 0; b < P.nvab; b++)
    if(P.d2_size[7 * b + 6] <= 0 || P.d2_size[7 * b + 6] > 4096)
      return fail(XT_EINVAL, f + ": bad p7 tile size");
  // the kernel keeps an exec slot in 30 bits of its LDS tile list
  for(size_t i = 0; i < 9 * P.noab; i++)
    if(P.d1_exec[i] >= (1 << 30)) return fail(XT_EINVAL, f + ": d1 exec slot out of range");
  for(size_t i = 0; i < 9 * P.nvab; i++)
    if(P.d2_exec[i] >= (1 << 30)) return fail(XT_EINVAL, f + ": d2 exec slot out of range");
  int* m   = meta;
  auto put = [&](const int* src, size_t n) {
    std::memcpy(m, src, n * sizeof(int));
    m += n;
  };
  put(P.d1_size, 7 * P.noab);
  put(P.d1_exec, 9 * P.noab);
  put(P.d2_size, 7 * P.nvab);
  put(P.d2_exec, 9 * P.nvab);
  put(P.s1_size, 6);
  put(P.s1_exec, 9);
  // K tile lists of the nine bipartitions g = 3*hs + ps: the enabled d1
  // tiles (h7b) with their exec slot for d1 equation 3*((ps+1)%3) + hs, then
  // the enabled d2 tiles (p7b) for d2 equation 3*ps + (hs+1)%3, packed end to
  // end except that the d2 tiles start on a chunk boundary
  // (ccsd_t_all_fused_cpu.hpp:173-270, 331-428)
  int* counts = m;
  m += xt::kKlistCountInts;
  for(int g = 0; g < 9; g++) {
    const int hs = g / 3, ps = g % 3, s1d = 3 * ((ps + 1) % 3) + hs, s2d = 3 * ps + (hs + 1) % 3;
    int       n = 0, nd1 = 0, kb = 0;
    auto      entry = [&](int fam, int ex, int K, int k0) {
      m[0] = fam;
      m[1] = ex;
      m[2] = K;
      m[3] = k0;
      m += 4;
    };
    for(size_t tb = 0; tb < P.noab; tb++) {
      const int ex = P.d1_exec[s1d + 9 * tb];
      if(ex < 0) continue;
      entry(0, ex, P.d1_size[3 + 7 * tb], kb);
      kb += P.d1_size[3 + 7 * tb];
      n++;
    }
    nd1 = n;
    for(size_t tb = 0; tb < P.nvab; tb++) {
      const int ex = P.d2_exec[s2d + 9 * tb];
      if(ex < 0) continue;
      if(n == nd1) kb = (kb + xt::kPackedKC - 1) / xt::kPackedKC * xt::kPackedKC;
      entry(1, ex, P.d2_size[6 + 7 * tb], kb);
      kb += P.d2_size[6 + 7 * tb];
      n++;
    }
    entry(0, 0, 0, kb);
    counts[g] = n | nd1 << 16;
  }
  *n_meta = (int) (m - meta);
  for(int i = 0; i < 6; i++) A.dim[i] = (int) P.d[i];
  A.d1_t2      = P.d1_t2;
  A.d1_v2      = P.d1_v2;
  A.d2_t2      = P.d2_t2;
  A.d2_v2      = P.d2_v2;
  A.s1_t1      = P.s1_t1;
  A.s1_v2      = P.s1_v2;
  A.meta       = nullptr;
  A.noab       = (int) P.noab;
  A.nvab       = (int) P.nvab;
  A.max_d1_t2  = (long) P.max_d1_t2;
  A.max_d1_v2  = (long) P.max_d1_v2;
  A.max_d2_t2  = (long) P.max_d2_t2;
  A.max_d2_v2  = (long) P.max_d2_v2;
  A.max_s1_t1  = (long) P.max_s1_t1;
  A.max_s1_v2  = (long) P.max_s1_v2;
  for(int i = 0; i < 6; i++) A.evl[i] = P.evl[i];
  A.partial    = P.partial;
  A.num_blocks = (long) P.num_blocks;
  return XT_OK;
}
}  // namespace

#define XT_PACKED_CALL                                                                              \
  PackedCall P = {num_blocks, {h1, h2, h3, p4, p5, p6}, d1_t2, d1_v2, d2_t2, d2_v2, s1_t1, s1_v2,  \
                  d1_size, d1_exec, d2_size, d2_exec, s1_size, s1_exec, noab, max_d1_t2,          \
                  max_d1_v2, nvab, max_d2_t2, max_d2_v2, max_s1_t1, max_s1_v2,                    \
                  {evl_h1, evl_h2, evl_h3, evl_p4, evl_p5, evl_p6}, partial}

int xt_fully_fused_ccsd_t_gpu(void* stream, size_t num_blocks, size_t h1, size_t h2, size_t h3,
                              size_t p4, size_t p5, size_t p6, const double* d1_t2,
                              const double* d1_v2, const double* d2_t2, const double* d2_v2,
                              const double* s1_t1, const double* s1_v2, const int* d1_size,
                              const int* d1_exec, const int* d2_size, const int* d2_exec,
                              const int* s1_size, const int* s1_exec, size_t noab,
                              size_t max_d1_t2, size_t max_d1_v2, size_t nvab, size_t max_d2_t2,
                              size_t max_d2_v2, size_t max_s1_t1, size_t max_s1_v2,
                              const double* evl_h1, const double* evl_h2, const double* evl_h3,
                              const double* evl_p4, const double* evl_p5, const double* evl_p6,
                              double* partial, void* done_copy) {
  XT_PACKED_CALL;
  // validate before touching any device state
  int            staged[kMetaInts];
  int            n_meta = 0;
  xt::PackedArgs A;
  int            rc = packed_prepare("xt_fully_fused_ccsd_t_gpu", P, staged, &n_meta, A);
  if(rc) return rc;
  PackedState& S = g_packed;
  int          dev;
  XT_HIP(hipGetDevice(&dev));
  if(S.device != dev) {
    if(S.device >= 0) return fail(XT_ESTATE, "xt_fully_fused_ccsd_t_gpu: device changed on thread");
    XT_HIP(hipHostMalloc((void**) &S.pinned, sizeof(int) * kMetaInts * kRing));
    XT_HIP(hipMalloc((void**) &S.dmeta, sizeof(int) * kMetaInts * kRing));
    for(auto& e: S.ev) XT_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    S.device = dev;
  }
  hipStream_t st   = (hipStream_t) stream;
  const int   slot = S.slot;
  S.slot           = (S.slot + 1) % kRing;
  XT_HIP(hipEventSynchronize(S.ev[slot]));  // previous use of this ring slot finished
  int* hm = S.pinned + slot * kMetaInts;
  std::memcpy(hm, staged, sizeof(int) * (size_t) n_meta);
  int* dm = S.dmeta + slot * kMetaInts;
  XT_HIP(hipMemcpyAsync(dm, hm, (size_t) n_meta * sizeof(int), hipMemcpyHostToDevice, st));
  if(done_copy) XT_HIP(hipEventRecord((hipEvent_t) done_copy, st));
  A.meta  = dm;
  int err = xt::launch_packed_fp64(A, st);
  if(err) return fail(XT_EHIP, std::string("packed launch: ") + hipGetErrorString((hipError_t) err));
  XT_HIP(hipEventRecord(S.ev[slot], st));
  return XT_OK;
}

int xt_fully_fused_ccsd_t_gpu_noalloc(
  void* stream, size_t num_blocks, size_t h1, size_t h2, size_t h3, size_t p4, size_t p5, size_t p6,
  const double* d1_t2, const double* d1_v2, const double* d2_t2, const double* d2_v2,
  const double* s1_t1, const double* s1_v2, const int* d1_size, const int* d1_exec,
  const int* d2_size, const int* d2_exec, const int* s1_size, const int* s1_exec, size_t noab,
  size_t max_d1_t2, size_t max_d1_v2, size_t nvab, size_t max_d2_t2, size_t max_d2_v2,
  size_t max_s1_t1, size_t max_s1_v2, const double* evl_h1, const double* evl_h2,
  const double* evl_h3, const double* evl_p4, const double* evl_p5, const double* evl_p6,
  double* partial, void* done_copy) {
  XT_PACKED_CALL;
  int            meta[kMetaInts];  // on the stack: nothing is allocated
  int            n_meta = 0;
  xt::PackedArgs A;
  int rc = packed_prepare("xt_fully_fused_ccsd_t_gpu_noalloc", P, meta, &n_meta, A);
  if(rc) return rc;
  if(n_meta > kMetaInts)  // cannot happen inside the noab/nvab limits checked above
    return fail(XT_EINVAL, "xt_fully_fused_ccsd_t_gpu_noalloc: metadata beyond the scratch size");
  // the metadata goes to the device in the copy kernels' arguments, so the
  // caller's arrays are free once the call returns; done_copy is recorded
  // after those copies, before the launch
  int err = xt::launch_packed_fp64_noalloc(A, meta, n_meta, stream, done_copy);
  if(err)
    return fail(XT_EHIP, std::string("xt_fully_fused_ccsd_t_gpu_noalloc: ") +
                           hipGetErrorString((hipError_t) err));
  return XT_OK;
}

}  // extern "C"
