// Stage 2 (fitter): per-stream launches with compensated (BMFR_QR_KAHAN) sums.
#include "fitter_body.h"

namespace bmfr {

hipError_t fitter_single_comp(const FitterParams &fp, bool tmp_f16, const FitterConfig &cfg, hipStream_t s) {
    return dispatch_single<true>(fp, tmp_f16, cfg, s);
}

}  // namespace bmfr
