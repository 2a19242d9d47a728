// Host emulation of the persistent kernels' work-queue claim loop
// (k_spatial_fp64 in xt_kernels.hip, k_spatial_mixed in xt_mixed.hip) over a
// queue layout the library built (xt_queue_layout).  Workgroup b serves queue
// b % 8 first, then the others in turn; each claim is one atomic increment
// of a queue's ticket counter, decoded with the kernels' own queue_entry_from
// (checked against queue_entry) and ticket_box (exachem_amd/csrc/xt_queue.h).
// The order in which workgroups
// claim is drawn at random (seeded), standing in for any interleaving of the
// atomics.  Checks: the layout's ticket ranges match task_ticket_count;
// every box of every selected task is claimed exactly once and no other;
// every workgroup leaves the loop, after at most (its tickets + 8) claims.
// Built and run by tests/test_queue_emulation.py.
//
//   queue_emu <layout file> <grid> <seed>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../exachem_amd/csrc/xt_queue.h"

namespace {
template<class T>
bool read_n(FILE* f, T* p, size_t n) {
  return std::fread(p, sizeof(T), n, f) == n;
}
int fail(const char* what, long long a = 0, long long b = 0) {
  std::printf("queue_emu: FAIL %s (%lld, %lld)\n", what, a, b);
  return 1;
}
}  // namespace

int main(int argc, char** argv) {
  if(argc != 4) return fail("usage: queue_emu <layout> <grid> <seed>");
  FILE* f = std::fopen(argv[1], "rb");
  if(!f) return fail("cannot open layout");
  const int      grid = std::atoi(argv[2]);
  const uint64_t seed = std::strtoull(argv[3], nullptr, 10);
  int64_t        hdr[2];
  if(!read_n(f, hdr, 2)) return fail("header");
  const long             nt = (long) hdr[0], ne = (long) hdr[1];
  std::vector<int>       qbeg(xt::kQueues + 1), qtask(ne), nbox(6 * nt), glog(6), sel(nt);
  std::vector<long long> qtick(ne), qtotal(xt::kQueues);
  bool ok = read_n(f, qbeg.data(), qbeg.size()) && read_n(f, qtask.data(), ne) &&
            read_n(f, qtick.data(), ne) && read_n(f, qtotal.data(), qtotal.size()) &&
            read_n(f, nbox.data(), nbox.size()) && read_n(f, glog.data(), 6) &&
            read_n(f, sel.data(), nt);
  std::fclose(f);
  if(!ok) return fail("truncated layout");
  if(grid < 1) return fail("grid", grid);

  // layout: contiguous ticket ranges of the kernels' size, every selected task once
  if(qbeg[0] != 0 || qbeg[xt::kQueues] != ne) return fail("qbeg bounds", qbeg[0], qbeg[8]);
  std::vector<int> listed(nt, 0);
  for(int q = 0; q < xt::kQueues; q++) {
    long long tk = 0;
    for(int e = qbeg[q]; e < qbeg[q + 1]; e++) {
      const int t = qtask[e];
      if(t < 0 || t >= nt || !sel[t]) return fail("entry names an unselected task", e, t);
      if(listed[t]++) return fail("task listed twice", t);
      if(qtick[e] != tk) return fail("ticket range not contiguous", e, qtick[e]);
      tk += xt::task_ticket_count(&nbox[6 * t], glog.data());
    }
    if(tk != qtotal[q]) return fail("queue total", q, qtotal[q]);
  }
  for(long t = 0; t < nt; t++)
    if(sel[t] && !listed[t]) return fail("selected task not in any queue", t);

  // box-claim counters of every task
  std::vector<long long> base(nt + 1, 0);
  for(long t = 0; t < nt; t++) {
    long long nb = 1;
    for(int p = 0; p < 6; p++) nb *= nbox[6 * t + p];
    base[t + 1] = base[t] + nb;
  }
  if(base[nt] > 400000000ll) return fail("too many boxes to emulate", base[nt]);
  std::vector<uint8_t> claimed(base[nt], 0);

  // the claim loop, one claim per step, workgroups drawn at random
  std::vector<unsigned long long> ctr(xt::kQueues, 0);
  std::vector<int>                qi(grid, 0);  // position in the workgroup's queue order
  std::vector<int>                hint(grid);   // entry of the workgroup's last ticket on its queue
  for(int b = 0; b < grid; b++) hint[b] = qbeg[b % xt::kQueues];
  std::vector<long long>          claims(grid, 0), got(grid, 0);
  std::vector<int>                live(grid);
  for(int b = 0; b < grid; b++) live[b] = b;
  std::mt19937_64 rng(seed);
  long long       boxes = 0;
  while(!live.empty()) {
    const size_t i = (size_t) (rng() % live.size());
    const int    b = live[i];
    const int    q = (b % xt::kQueues + qi[b]) % xt::kQueues;
    const unsigned long long t = ctr[q]++;  // atomicAdd(&R.qctr[q], 1)
    claims[b]++;
    if((long long) t >= qtotal[q]) {  // queue exhausted: next queue, or leave
      if(++qi[b] == xt::kQueues) {
        live[i] = live.back();
        live.pop_back();
      }
      else hint[b] = qbeg[(b % xt::kQueues + qi[b]) % xt::kQueues];
      continue;
    }
    // the kernels search forward from the workgroup's previous entry; that
    // must find the entry a search over the whole queue finds
    const int e = xt::queue_entry_from(qbeg.data(), qtick.data(), q, (long long) t, hint[b]);
    if(e != xt::queue_entry(qbeg.data(), qtick.data(), q, (long long) t))
      return fail("queue_entry_from", e, (long long) t);
    hint[b] = e;
    const int       task  = qtask[e];
    const long long local = (long long) t - qtick[e];
    if(e < qbeg[q] || e >= qbeg[q + 1] || local < 0) return fail("queue_entry", e, local);
    int sub[6];
    if(!xt::ticket_box(local, &nbox[6 * task], glog.data(), sub)) continue;
    long long idx = 0;
    for(int p = 5; p >= 0; p--) {
      if(sub[p] < 0 || sub[p] >= nbox[6 * task + p]) return fail("box outside grid", task, p);
      idx = idx * nbox[6 * task + p] + sub[p];
    }
    if(claimed[base[task] + idx]++) return fail("box claimed twice", task, idx);
    got[b]++;
    boxes++;
  }
  long long want = 0;
  for(long t = 0; t < nt; t++)
    for(long long k = base[t]; k < base[t + 1]; k++) {
      if(claimed[k] != (sel[t] ? 1 : 0)) return fail("box claim count", t, k - base[t]);
      want += sel[t];
    }
  // each workgroup stops after its tickets plus one exhausted claim per queue
  for(int b = 0; b < grid; b++)
    if(claims[b] < xt::kQueues) return fail("workgroup left early", b, claims[b]);
  long long total_t = 0;
  for(long long x: qtotal) total_t += x;
  long long total_c = 0;
  for(long long x: claims) total_c += x;
  if(total_c != total_t + (long long) grid * xt::kQueues) return fail("claims", total_c, total_t);
  long long busiest = 0;
  for(long long x: got) busiest = x > busiest ? x : busiest;
  std::printf("queue_emu: ok tasks %ld boxes %lld tickets %lld grid %d busiest %lld\n", nt, boxes,
              total_t, grid, busiest);
  return want == boxes ? 0 : fail("box total", want, boxes);
}
