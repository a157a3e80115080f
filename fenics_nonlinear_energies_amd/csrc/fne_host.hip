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

// Greedy distance-2 colouring (see include/fne_hip.h).  forbid[c] holds the last row for which
// colour c was found in the distance-2 neighbourhood, so the marks need no clearing between rows.
extern "C" int fne_color_distance2(int32_t n, const int32_t* indptr, const int32_t* indices, int32_t* colors) {
  if (n < 0 || (n > 0 && (!indptr || !indices || !colors))) {
    fne_set_error("fne_color_distance2: invalid arguments");
    return FNE_EINVAL;
  }
  if (n == 0) return 0;
  if (indptr[0] != 0) {
    fne_set_error("fne_color_distance2: indptr[0] must be 0");
    return FNE_EINVAL;
  }
  for (int32_t i = 0; i < n; ++i) {
    if (indptr[i + 1] < indptr[i]) {
      fne_set_error("fne_color_distance2: indptr must be non-decreasing");
      return FNE_EINVAL;
    }
  }
  const int64_t nnz = indptr[n];
  for (int64_t k = 0; k < nnz; ++k) {
    if (indices[k] < 0 || indices[k] >= n) {
      fne_set_error("fne_color_distance2: column index out of range");
      return FNE_EINVAL;
    }
  }
  for (int32_t i = 0; i < n; ++i) colors[i] = -1;
  std::vector<int32_t> forbid;
  int32_t n_colors = 0;
  for (int32_t v = 0; v < n; ++v) {
    for (int32_t k = indptr[v]; k < indptr[v + 1]; ++k) {
      const int32_t u = indices[k];
      if (colors[u] >= 0) forbid[size_t(colors[u])] = v;
      for (int32_t kk = indptr[u]; kk < indptr[u + 1]; ++kk) {
        const int32_t w = indices[kk];
        if (colors[w] >= 0) forbid[size_t(colors[w])] = v;
      }
    }
    int32_t c = 0;
    while (c < n_colors && forbid[size_t(c)] == v) ++c;
    if (c == n_colors) {
      forbid.push_back(-1);
      ++n_colors;
    }
    colors[v] = c;
  }
  return n_colors;
}
