// Stage 2 (fitter): batched persistent launches with compensated (BMFR_QR_KAHAN) sums.
#include "fitter_body.h"

namespace bmfr {

hipError_t fitter_batch_comp(const BatchLaunch &L, int total, bool tmp_f16, const FitterConfig &cfg, int grid_cap, hipStream_t s) {
    return dispatch_batch<true>(L, total, tmp_f16, cfg, grid_cap, s);
}

}  // namespace bmfr
