// Stage 2 (fitter) for configured feature sets of 9..16 features.
#include "fitter_body.h"

namespace bmfr {

hipError_t fitter_set_hi(const FitterParams &fp, bool tmp_f16, int features, hipStream_t s) {
    switch (features) {
        case 9: return launch_set<9>(fp, tmp_f16, s);
        case 10: return launch_set<10>(fp, tmp_f16, s);
        case 11: return launch_set<11>(fp, tmp_f16, s);
        case 12: return launch_set<12>(fp, tmp_f16, s);
        case 13: return launch_set<13>(fp, tmp_f16, s);
        case 14: return launch_set<14>(fp, tmp_f16, s);
        case 15: return launch_set<15>(fp, tmp_f16, s);
        case 16: return launch_set<16>(fp, tmp_f16, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace bmfr
