// Stage 2 (fitter): batched persistent launches with plain FP32 sums (the fast body).
#include "fitter_body.h"

namespace bmfr {

hipError_t fitter_batch_plain(const BatchLaunch &L, int total, bool tmp_f16, const FitterConfig &cfg, int grid_cap, hipStream_t s) {
    return dispatch_batch<false>(L, total, tmp_f16, cfg, grid_cap, s);
}

}  // namespace bmfr
