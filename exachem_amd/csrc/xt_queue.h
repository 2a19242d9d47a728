// Work-queue tickets of the persistent kernels: which queue entry owns a
// ticket, and which box of the task's box grid a ticket names.  Shared by the
// kernels (xt_kernels.hip, xt_mixed.hip), the host planner (xt_host.cpp) and
// the host emulation of the claim loop (tests/host/queue_emu.cpp), so the
// emulation runs the kernels' own decode.
#pragma once
#include "xt_device.h"

#if defined(__HIPCC__)
#define XT_HD __host__ __device__
#else
#define XT_HD
#endif

namespace xt {

// Tickets of a task whose box grid is nbox[0..5] boxes along (h1,h2,h3,p4,p5,p6)
// under group shape glog: one group of 2^kGroupLog2 tickets per cell of the
// grid cut into 2^glog[p] boxes along each position.
XT_HD inline long long task_ticket_count(const int* nbox, const int* glog) {
  long long n = 1ll << kGroupLog2;
  for(int p = 0; p < 6; p++) n *= (nbox[p] + (1 << glog[p]) - 1) >> glog[p];
  return n;
}

// Entry of queue q that owns ticket t: the last entry e of qbeg[q]..qbeg[q+1]-1
// with qtick[e] <= t (qtick ascends within a queue; the caller checks
// t < qtotal[q], which makes the queue non-empty).
XT_HD inline int queue_entry(const int* qbeg, const long long* qtick, int q, long long t) {
  int lo = qbeg[q], hi = qbeg[q + 1] - 1;
  while(lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if(qtick[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// The same entry, searched from entry `from` of queue q, which must be at or
// before the answer (the entry of the caller's previous ticket on this queue:
// a workgroup's tickets on one queue only grow): two steps forward, then a
// binary search over the rest.  A workgroup's next ticket is usually in the
// same entry or the next, so this reads one or two qtick values instead of
// the dozen dependent reads of a search over the whole queue.
XT_HD inline int queue_entry_from(const int* qbeg, const long long* qtick, int q, long long t,
                                  int from) {
  int       lo   = from < qbeg[q] ? qbeg[q] : from;
  const int last = qbeg[q + 1] - 1;
  for(int s = 0; s < 2; s++) {
    if(lo >= last || qtick[lo + 1] > t) return lo;
    lo++;
  }
  int hi = last;
  while(lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if(qtick[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Box named by a task's local ticket: the high bits pick the group (mixed
// radix over the group grid, h1 fastest), the low kGroupLog2 bits the box
// inside it (glog[0] bits for h1, then glog[1] for h2, ...).  Returns false
// for a ticket whose box lies outside the grid (groups overhang the grid edge).
// 32-bit arithmetic: the host runs no task of 2^31 or more tickets.
XT_HD inline bool ticket_box(long long local, const int* nbox, const int* glog, int* sub) {
  unsigned  box   = (unsigned) (local >> kGroupLog2);
  const int inb   = (int) (local & ((1 << kGroupLog2) - 1));
  bool      ok    = true;
  int       shift = 0;
#pragma unroll
  for(int p = 0; p < 6; p++) {
    const int      gl = glog[p];
    const unsigned nb = (unsigned) ((nbox[p] + (1 << gl) - 1) >> gl);
    sub[p]            = ((int) (box % nb) << gl) + ((inb >> shift) & ((1 << gl) - 1));
    box /= nb;
    shift += gl;
    ok = ok && sub[p] < nbox[p];
  }
  return ok;
}

}  // namespace xt
