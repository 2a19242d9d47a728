// Host planning of the device-side per-task gather: which blocks the
// reference's gathers copy for one task, in which order, with which
// permutation, plus the metadata the fused launcher reads.  Restates
//   ccsd_t_data_s1_new  ccsd_t_all_fused_singles.hpp:34-391
//   ccsd_t_data_d1_new  ccsd_t_all_fused_doubles1.hpp:37-307
//   ccsd_t_data_d2_new  ccsd_t_all_fused_doubles2.hpp:37-360
// and the buffer sizing of ccsd_t_fused_driver.hpp:176-195 (tilesize_opt).
// The element values are formed on the device (xt_gather.hip).
#include <algorithm>
#include <array>
#include <vector>

#include "xt_device.h"
#include "xt_internal.h"

namespace xt {
namespace {

using Tuple = std::array<long, 6>;  // (p4b, p5b, p6b, h1b, h2b, h3b)

// Later duplicates of a tuple are dropped (set to all zeros), as the gathers
// do before enumerating slots.
void drop_duplicates(Tuple (&a)[9]) {
  for(int i = 0; i < 8; i++)
    if(a[i][0] != 0)
      for(int j = i + 1; j < 9; j++)
        if(a[i] == a[j]) a[j] = Tuple{0, 0, 0, 0, 0, 0};
}

struct Tiles {
  const GatherTiling& T;
  int  spin(long b) const { return T.k_spin[b]; }
  int  range(long b) const { return T.k_range[b]; }
  bool all_nonempty(const Tuple& t) const {
    for(long b: t)
      if(T.k_range[b] <= 0) return false;
    return true;
  }
  // tuple-wide tests shared by the three gathers
  bool spin_ok(const Tuple& t, bool restricted) const {
    const int sp = spin(t[0]) + spin(t[1]) + spin(t[2]), sh = spin(t[3]) + spin(t[4]) + spin(t[5]);
    if(restricted && sp + sh == 12) return false;
    return sp == sh;
  }
};

GatherBlock block(int buf, int kind, long dst, const Tiles& tl, const long* tiles, int nd,
                  const int* perm, double alpha) {
  GatherBlock B = {};
  B.buf         = buf;
  B.kind        = kind;
  B.dst         = dst;
  B.nd          = nd;
  B.alpha       = alpha;
  B.n           = 1;
  for(int d = 0; d < nd; d++) {
    B.size[d]  = tl.range(tiles[d]);
    B.start[d] = tl.T.spatial_off[tiles[d]];
    B.spin[d]  = tl.spin(tiles[d]) == 2 ? 1 : 0;
    B.perm[d]  = perm[d];
    B.n *= B.size[d];
  }
  return B;
}

}  // namespace

void packed_sizes(const GatherTiling& T, long out[8]) {
  long max_p = 0, max_h = 0, noa = 0, nva = 0;
  for(int b = 0; b < T.noab; b++) {
    max_h = std::max<long>(max_h, T.k_range[b]);
    noa += T.k_spin[b] == 1;
  }
  for(int b = T.noab; b < T.noab + T.nvab; b++) {
    max_p = std::max<long>(max_p, T.k_range[b]);
    nva += T.k_spin[b] == 1;
  }
  const long kd1 = 9 * noa, kd2 = 9 * nva;
  out[0] = 9 * max_p * max_h;                    // s1_t1
  out[1] = 9 * (max_p * max_p) * (max_h * max_h);  // s1_v2
  out[2] = kd1 * (max_p * max_p) * (max_h * max_h);  // d1_t2
  out[3] = kd1 * max_p * (max_h * max_h * max_h);  // d1_v2
  out[4] = kd2 * (max_p * max_p) * (max_h * max_h);  // d2_t2
  out[5] = kd2 * (max_p * max_p * max_p) * max_h;  // d2_v2
  out[6] = kd1;                                  // max d1 kernels per task
  out[7] = kd2;                                  // max d2 kernels per task
}

int plan_gather(const GatherTiling& T, bool restricted, const long* t, std::vector<GatherBlock>& out,
                int* s1_size, int* s1_exec, int* d1_size, int* d1_exec, int* d2_size, int* d2_exec,
                int* counts) {
  const Tiles tl{T};
  const long  noab = T.noab, nvab = T.nvab;
  const long  h1 = t[0], h2 = t[1], h3 = t[2], p4 = t[3], p5 = t[4], p6 = t[5];
  const Tuple ref = {p4, p5, p6, h1, h2, h3};
  long        sz[8];
  packed_sizes(T, sz);
  out.clear();
  std::fill(s1_exec, s1_exec + 9, -1);
  std::fill(d1_exec, d1_exec + 9 * noab, -1);
  std::fill(d2_exec, d2_exec + 9 * nvab, -1);

  // ---- s1 (singles.hpp:55-212, 304-314)
  {
    Tuple a[9] = {{p4, p5, p6, h1, h2, h3}, {p4, p5, p6, h2, h1, h3}, {p4, p5, p6, h3, h1, h2},
                  {p5, p4, p6, h1, h2, h3}, {p5, p4, p6, h2, h1, h3}, {p5, p4, p6, h3, h1, h2},
                  {p6, p4, p5, h1, h2, h3}, {p6, p4, p5, h2, h1, h3}, {p6, p4, p5, h3, h1, h2}};
    drop_duplicates(a);
    for(int k = 0; k < 6; k++) s1_size[k] = tl.range(t[k]);
    const long dima = sz[0] / 9, dimb = sz[1] / 9;
    int        idx = 0;
    for(int s = 0; s < 9; s++) {
      const Tuple& x = a[s];
      const long   P4 = x[0], P5 = x[1], P6 = x[2], H1 = x[3], H2 = x[4], H3 = x[5];
      if(!(P5 <= P6 && H2 <= H3 && P4 != 0)) continue;
      if(!tl.spin_ok(x, restricted) || tl.spin(P4) != tl.spin(H1) || !tl.all_nonempty(x)) continue;
      const Tuple pat[9] = {{P4, P5, P6, H1, H2, H3}, {P4, P5, P6, H2, H1, H3}, {P4, P5, P6, H2, H3, H1},
                            {P5, P4, P6, H1, H2, H3}, {P5, P4, P6, H2, H1, H3}, {P5, P4, P6, H2, H3, H1},
                            {P5, P6, P4, H1, H2, H3}, {P5, P6, P4, H2, H1, H3}, {P5, P6, P4, H2, H3, H1}};
      for(int k = 0; k < 9; k++)
        if(ref == pat[k]) s1_exec[k] = idx;
      const long bt1[2] = {P4, H1};
      const int  pm1[2] = {1, 0};
      out.push_back(block(4, kGatherT1, idx * dima, tl, bt1, 2, pm1, 1.0));
      const long bt2[4] = {H3, H2, P6, P5};
      const int  pm2[4] = {3, 2, 1, 0};
      out.push_back(block(5, kGatherIJAB, idx * dimb, tl, bt2, 4, pm2, 1.0));
      idx++;
    }
    counts[0] = idx;
  }

  // ---- d1 (doubles1.hpp:50-290), summed over h7
  {
    Tuple a[9] = {{p4, p5, p6, h1, h2, h3}, {p4, p5, p6, h2, h1, h3}, {p4, p5, p6, h3, h1, h2},
                  {p5, p6, p4, h1, h2, h3}, {p5, p6, p4, h2, h1, h3}, {p5, p6, p4, h3, h1, h2},
                  {p4, p6, p5, h1, h2, h3}, {p4, p6, p5, h2, h1, h3}, {p4, p6, p5, h3, h1, h2}};
    drop_duplicates(a);
    for(long h7 = 0; h7 < noab; h7++) {
      int* z = d1_size + 7 * h7;
      z[0] = tl.range(h1), z[1] = tl.range(h2), z[2] = tl.range(h3), z[3] = tl.range(h7);
      z[4] = tl.range(p4), z[5] = tl.range(p5), z[6] = tl.range(p6);
    }
    const long dima = sz[2] / sz[6], dimb = sz[3] / sz[6];
    int        idx = 0;
    for(int s = 0; s < 9; s++) {
      const Tuple& x = a[s];
      const long   P4 = x[0], P5 = x[1], P6 = x[2], H1 = x[3], H2 = x[4], H3 = x[5];
      if(!(P4 <= P5 && H2 <= H3 && P4 != 0) || !tl.spin_ok(x, restricted)) continue;
      const Tuple pat[9] = {{P4, P5, P6, H1, H2, H3}, {P4, P5, P6, H2, H1, H3}, {P4, P5, P6, H2, H3, H1},
                            {P6, P4, P5, H1, H2, H3}, {P6, P4, P5, H2, H1, H3}, {P6, P4, P5, H2, H3, H1},
                            {P4, P6, P5, H1, H2, H3}, {P4, P6, P5, H2, H1, H3}, {P4, P6, P5, H2, H3, H1}};
      for(long h7 = 0; h7 < noab; h7++) {
        if(tl.spin(P4) + tl.spin(P5) != tl.spin(H1) + tl.spin(h7)) continue;
        if(!tl.all_nonempty(x) || !(h7 <= P6)) continue;
        for(int k = 0; k < 9; k++)
          if(ref == pat[k]) d1_exec[k + 9 * h7] = idx;
        if(h7 < H1) {  // doubles1.hpp:218-232: t2(p4,p5,h7,h1), perm {3,1,0,2}, -1
          const long bt[4] = {P4, P5, h7, H1};
          const int  pm[4] = {3, 1, 0, 2};
          out.push_back(block(0, kGatherT2, idx * dima, tl, bt, 4, pm, -1.0));
        }
        else {  // doubles1.hpp:233-247: t2(p4,p5,h1,h7), perm {2,1,0,3}
          const long bt[4] = {P4, P5, H1, h7};
          const int  pm[4] = {2, 1, 0, 3};
          out.push_back(block(0, kGatherT2, idx * dima, tl, bt, 4, pm, 1.0));
        }
        // doubles1.hpp:278-290: v2ijka(h2,h3,h7,p6), perm {2,3,0,1}
        const long bt[4] = {H2, H3, h7, P6};
        const int  pm[4] = {2, 3, 0, 1};
        out.push_back(block(1, kGatherIJKA, idx * dimb, tl, bt, 4, pm, 1.0));
        idx++;
      }
    }
    counts[1] = idx;
  }

  // ---- d2 (doubles2.hpp:50-330), summed over p7
  {
    Tuple a[9] = {{p4, p5, p6, h1, h2, h3}, {p4, p5, p6, h2, h3, h1}, {p4, p5, p6, h1, h3, h2},
                  {p5, p4, p6, h1, h2, h3}, {p5, p4, p6, h2, h3, h1}, {p5, p4, p6, h1, h3, h2},
                  {p6, p4, p5, h1, h2, h3}, {p6, p4, p5, h2, h3, h1}, {p6, p4, p5, h1, h3, h2}};
    drop_duplicates(a);
    for(long p7 = noab; p7 < noab + nvab; p7++) {
      int* z = d2_size + 7 * (p7 - noab);
      z[0] = tl.range(h1), z[1] = tl.range(h2), z[2] = tl.range(h3), z[3] = tl.range(p4);
      z[4] = tl.range(p5), z[5] = tl.range(p6), z[6] = tl.range(p7);
    }
    const long dima = sz[4] / sz[7], dimb = sz[5] / sz[7];
    int        idx = 0;
    for(int s = 0; s < 9; s++) {
      const Tuple& x = a[s];
      const long   P4 = x[0], P5 = x[1], P6 = x[2], H1 = x[3], H2 = x[4], H3 = x[5];
      if(!(P5 <= P6 && H1 <= H2 && P4 != 0) || !tl.spin_ok(x, restricted)) continue;
      const Tuple pat[9] = {{P4, P5, P6, H1, H2, H3}, {P4, P5, P6, H3, H1, H2}, {P4, P5, P6, H1, H3, H2},
                            {P5, P4, P6, H1, H2, H3}, {P5, P4, P6, H3, H1, H2}, {P5, P4, P6, H1, H3, H2},
                            {P5, P6, P4, H1, H2, H3}, {P5, P6, P4, H3, H1, H2}, {P5, P6, P4, H1, H3, H2}};
      for(long p7 = noab; p7 < noab + nvab; p7++) {
        if(tl.spin(P4) + tl.spin(p7) != tl.spin(H1) + tl.spin(H2)) continue;
        if(!tl.all_nonempty(x) || !(H3 <= p7)) continue;
        for(int k = 0; k < 9; k++)
          if(ref == pat[k]) d2_exec[k + 9 * (p7 - noab)] = idx;
        if(p7 < P4) {  // doubles2.hpp:211-225: t2(p7,p4,h1,h2), perm {3,2,1,0}, -1
          const long bt[4] = {p7, P4, H1, H2};
          const int  pm[4] = {3, 2, 1, 0};
          out.push_back(block(2, kGatherT2, idx * dima, tl, bt, 4, pm, -1.0));
        }
        else {  // doubles2.hpp:226-239: t2(p4,p7,h1,h2), perm {3,2,0,1}
          const long bt[4] = {P4, p7, H1, H2};
          const int  pm[4] = {3, 2, 0, 1};
          out.push_back(block(2, kGatherT2, idx * dima, tl, bt, 4, pm, 1.0));
        }
        // doubles2.hpp:335-345: v2iabc(h3,p7,p5,p6), perm {2,3,0,1}
        const long bt[4] = {H3, p7, P5, P6};
        const int  pm[4] = {2, 3, 0, 1};
        out.push_back(block(3, kGatherIABC, idx * dimb, tl, bt, 4, pm, 1.0));
        idx++;
      }
    }
    counts[2] = idx;
  }
  return 0;
}

}  // namespace xt
