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
  int      dim[6];      // tile extents
  int      off[6];      // first spatial orbital of each tile (occ or virt range)
  int      nsub[6];     // ceil(dim/4): 4-wide sub-tiles per position
  int      n_sub;       // product of nsub
  int      unit_begin;  // first work unit of this task
  int      n_units;     // work units of this task
  int      subs_per_unit;
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

// HBM-resident closed-shell inputs in kernel layouts.
struct SpatialDev {
  int           O, V;
  const double* eo;    // [O]
  const double* ev;    // [V]
  const double* t1;    // [V][O]          t_i^a
  const double* t2h;   // [V][V][O][O]    t_ij^ab            (j contiguous; d1 operand)
  const double* t2p;   // [O][O][V][V]    t2p[i][j][a][b] = t_ij^ab (b contiguous; d2 operand)
  const double* ovk;   // [O][V][O][O]    (ia|jk)            (k contiguous; d1 operand)
  const double* xv;    // [V][V][O][V]    xv[c][b][i][a] = (ac|bi) (a contiguous; d2 operand)
  const double* vvoo;  // [V][V][O][O]    (ai|bj)            (s1 operand)
};

struct RunDev {
  const TaskDev* tasks;
  const int*     unit_task;   // [n_units] task of each unit
  const int*     order;       // [n_units] queue order (unit ids)
  int            n_units;
  int            n_tasks;
  double*        unit_e;      // [2*n_units] per-unit (E4,E5) partials, no factor
  unsigned int*  queue;       // device work counter
  unsigned long long* task_t0;  // optional per-task first/last timestamps
  unsigned long long* task_t1;
};

// Kernel launchers (defined in xt_kernels.hip).  Return hipError_t as int.
int launch_relayout(const double* t2_in, const double* ovoo_in, const double* vvvo_in, int O,
                    int V, double* t2h, double* t2p, double* ovk, double* xv, void* stream);
int launch_spatial_fp64(const SpatialDev& S, const RunDev& R, int grid, void* stream);
int launch_reduce(const TaskDev* tasks, int n_tasks, const double* unit_e, double* task_e,
                  double* e45, void* stream);

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

}  // namespace xt
