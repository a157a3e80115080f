// Host-side plan-construction helpers (plain C++, no device code).
#include <cstring>
#include <vector>

#include "fne_common.h"

extern "C" int fne_color_patch_elements(int32_t n_patch, const int64_t* pelem_ptr, const int64_t* rows,
                                        int32_t n_loc, int64_t n_rows, int32_t* colors) {
  if (n_patch < 0 || n_loc <= 0 || n_rows < 0 || (n_patch > 0 && (!pelem_ptr || !rows || !colors))) {
    fne_set_error("fne_color_patch_elements: invalid arguments");
    return FNE_EINVAL;
  }
  // rows of different patches are disjoint, so one mask per row serves all patches
  std::vector<uint64_t> mask(size_t(n_rows), 0ull);
  int max_colors = 0;
  for (int32_t p = 0; p < n_patch; ++p) {
    int patch_colors = 0;
    for (int64_t i = pelem_ptr[p]; i < pelem_ptr[p + 1]; ++i) {
      uint64_t used = 0ull;
      for (int a = 0; a < n_loc; ++a) {
        const int64_t r = rows[i * n_loc + a];
        if (r >= 0) {
          if (r >= n_rows) {
            fne_set_error("fne_color_patch_elements: row index out of range");
            return FNE_EINVAL;
          }
          used |= mask[size_t(r)];
        }
      }
      if (used == ~0ull) {
        fne_set_error("fne_color_patch_elements: more than 64 colors needed in one patch");
        return FNE_EINVAL;
      }
      const int c = __builtin_ctzll(~used);
      colors[i] = c;
      if (c + 1 > patch_colors) patch_colors = c + 1;
      for (int a = 0; a < n_loc; ++a) {
        const int64_t r = rows[i * n_loc + a];
        if (r >= 0) mask[size_t(r)] |= (1ull << c);
      }
    }
    if (patch_colors > max_colors) max_colors = patch_colors;
  }
  return max_colors;
}
