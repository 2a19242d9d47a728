// Host-side helpers shared by the library's translation units.
#pragma once
#include <string>

#include <vector>

namespace xt {
// Records msg as the calling thread's last error (xt_last_error) and returns code.
int set_error(int code, const std::string& msg);

struct GatherBlock;
// Spin-orbital tiling as the gathers see it: tiles [0, noab) occupied,
// [noab, noab + nvab) virtual; k_spin 1 (alpha) or 2 (beta); spatial_off the
// tile's first spatial orbital in its space.
struct GatherTiling {
  int        noab, nvab;
  const int* k_range;
  const int* k_spin;
  const int* spatial_off;
};
// Packed buffer sizes of the fused driver: s1_t1, s1_v2, d1_t2, d1_v2, d2_t2,
// d2_v2 (doubles), then the maximum d1 and d2 slots per task.
void packed_sizes(const GatherTiling& T, long out[8]);
// Blocks and launcher metadata of the gathers for task t = (h1b, h2b, h3b,
// p4b, p5b, p6b).  exec arrays are set to -1 where no slot runs; counts =
// enabled s1, d1, d2 slots.
int plan_gather(const GatherTiling& T, bool restricted, const long* t, std::vector<GatherBlock>& out,
                int* s1_size, int* s1_exec, int* d1_size, int* d1_exec, int* d2_size, int* d2_exec,
                int* counts);
}  // namespace xt
