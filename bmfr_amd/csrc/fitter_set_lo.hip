// Stage 2 (fitter) for configured feature sets of 1..8 features.
#include "fitter_body.h"

namespace bmfr {

hipError_t fitter_set_lo(const FitterParams &fp, bool tmp_f16, int features, hipStream_t s) {
    switch (features) {
        case 1: return launch_set<1>(fp, tmp_f16, s);
        case 2: return launch_set<2>(fp, tmp_f16, s);
        case 3: return launch_set<3>(fp, tmp_f16, s);
        case 4: return launch_set<4>(fp, tmp_f16, s);
        case 5: return launch_set<5>(fp, tmp_f16, s);
        case 6: return launch_set<6>(fp, tmp_f16, s);
        case 7: return launch_set<7>(fp, tmp_f16, s);
        case 8: return launch_set<8>(fp, tmp_f16, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace bmfr
