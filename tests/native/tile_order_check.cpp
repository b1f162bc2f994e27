// Host build of the control-tree tile order (momms_amd/csrc/tile_order.hpp,
// the code the tree-ordered GEMM kernel runs). Reads lines
//   TM TN nsteps kind_1 bsz_1 ... kind_n bsz_n     (innermost step first)
// and prints, per line, the tiles (tm,tn) of logical ids 0 .. TM*TN-1 in
// order, as "tm,tn" separated by spaces ("ERR" when the steps are refused).
#include <cstdio>
#include <vector>

#include "tile_order.hpp"

using namespace momms;

int main() {
  long long TM, TN, ns;
  while (std::scanf("%lld %lld %lld", &TM, &TN, &ns) == 3) {
    std::vector<int32_t> kinds(ns > 0 ? ns : 1);
    std::vector<int64_t> bszs(ns > 0 ? ns : 1);
    for (long long i = 0; i < ns; ++i) {
      long long kd, b;
      if (std::scanf("%lld %lld", &kd, &b) != 2) return 2;
      kinds[i] = (int32_t)kd;
      bszs[i] = (int64_t)b;
    }
    TileOrder o;
    if (!tile_order_from_steps(kinds.data(), bszs.data(), ns, 128, 128, o)) {
      std::printf("ERR\n");
      continue;
    }
    for (long long t = 0; t < TM * TN; ++t) {
      int64_t tm, tn;
      tile_order_decode(t, TM, TN, o, tm, tn);
      std::printf("%s%lld,%lld", t ? " " : "", (long long)tm, (long long)tn);
    }
    std::printf("\n");
  }
  return 0;
}
