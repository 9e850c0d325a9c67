// Stage 2 (fitter): per-stream launches with plain FP32 sums (the fast body).
#include "fitter_body.h"

namespace bmfr {

hipError_t fitter_single_plain(const FitterParams &fp, bool tmp_f16, const FitterConfig &cfg, hipStream_t s) {
    return dispatch_single<false>(fp, tmp_f16, cfg, s);
}

}  // namespace bmfr
