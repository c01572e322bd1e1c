// Host check of tri_tile_coords (eigen_amd/csrc/common.hpp): for every grid
// of nt x nt tiles up to 300, the indices 0 .. nt(nt+1)/2 - 1 map one to one
// onto the tiles (r, c) with c <= r, and each group of 8 tile-rows is walked
// column by column.  Built and run by tests/test_capi_cpu.py.
#include <cstdio>
#include <vector>

#include "../eigen_amd/csrc/common.hpp"

int main() {
  for (int nt = 1; nt <= 300; ++nt) {
    const int total = nt * (nt + 1) / 2;
    std::vector<int> seen(nt * nt, 0);
    int prev_group = 0;
    for (int L = 0; L < total; ++L) {
      int r = -1, c = -1;
      eh::tri_tile_coords(L, nt, 8, r, c);
      if (r < 0 || r >= nt || c < 0 || c > r) {
        std::printf("FAIL nt=%d L=%d -> (%d, %d) outside the triangle\n", nt, L, r, c);
        return 1;
      }
      if (seen[r * nt + c]++) {
        std::printf("FAIL nt=%d L=%d -> (%d, %d) twice\n", nt, L, r, c);
        return 1;
      }
      if (r / 8 < prev_group) {
        std::printf("FAIL nt=%d L=%d: group order\n", nt, L);
        return 1;
      }
      prev_group = r / 8;
    }
  }
  std::printf("ALL PASSED\n");
  return 0;
}
