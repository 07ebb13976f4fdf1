// palx_launch.hip -- instantiates the row / column pass kernels of one
// precision.  Built twice: as is (float, launch_*_f32) and with
// -DPALX_REAL_IS_DOUBLE (double, launch_*_f64).
#include "palx_kernels.h"

namespace palx {

// every measured default step variant names a variant compiled for its
// length (in this file's precision: it exists and its block fits the LDS)
#if defined(PALX_REAL_IS_DOUBLE)
constexpr int kDefaultsCsz = 16;
#else
constexpr int kDefaultsCsz = 8;
#endif
constexpr bool default_variants_compiled() {
  for (const DefaultVariants& d : kDefaultStepVariants) {
    const int r = kDefaultsCsz == 8 ? d.f32_row : d.f64_row, c = kDefaultsCsz == 8 ? d.f32_col : d.f64_col;
    if (r >= 0 && !variant_fits(d.M, r, kDefaultsCsz)) return false;
    if (c >= 0 && !variant_fits(d.M, c, kDefaultsCsz)) return false;
  }
  return true;
}
static_assert(default_variants_compiled(), "a default step variant is not compiled for its length");

#if defined(PALX_REAL_IS_DOUBLE)
hipError_t launch_row_f64(int M, int V, int mode, const RowArgs& one, const RowArgs* tab, int nmem,
                          const void* tw, hipStream_t s) {
  return launch_row_any<double>(M, V, mode, one, tab, nmem, tw, s);
}
hipError_t launch_col_f64(int M, int V, int mode, int stage, const ColArgs& one, const ColArgs* tab,
                          int nmem, const void* tw, hipStream_t s) {
  return launch_col_any<double>(M, V, mode, stage, one, tab, nmem, tw, s);
}
#else
hipError_t launch_row_f32(int M, int V, int mode, const RowArgs& one, const RowArgs* tab, int nmem,
                          const void* tw, hipStream_t s) {
  return launch_row_any<float>(M, V, mode, one, tab, nmem, tw, s);
}
hipError_t launch_col_f32(int M, int V, int mode, int stage, const ColArgs& one, const ColArgs* tab,
                          int nmem, const void* tw, hipStream_t s) {
  return launch_col_any<float>(M, V, mode, stage, one, tab, nmem, tw, s);
}
int n_pass_variants(int M) {
  switch (M) {
#define PALX_CASE(L_) case L_: return n_variants(L_);
    PALX_FOR_EACH_LENGTH(PALX_CASE)
    PALX_FOR_EACH_LENGTH_F32(PALX_CASE)
#undef PALX_CASE
  }
  return 0;
}
#endif

}  // namespace palx
