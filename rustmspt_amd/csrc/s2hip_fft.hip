// FP32 real-FFT autocorrelation path (placeholder until the hand-written
// transform lands: every entry point reports S2H_ERR_UNSUPPORTED).
#include "s2hip_internal.h"

namespace s2h {
struct FftPlan {
    int unused;
};
void fft_plan_release(s2h_ctx *, FftPlan *p) { delete p; }
}  // namespace s2h

using namespace s2h;

extern "C" {
s2h_status s2h_fft32_counts(s2h_ctx *ctx, s2h_vol *vol, uint32_t, double, float *, uint64_t) {
    if (!ctx || !vol) return fail(S2H_ERR_ARG, "null argument");
    return fail(S2H_ERR_UNSUPPORTED, "FP32 FFT path not built yet");
}
s2h_status s2h_fft32_s2(s2h_ctx *ctx, s2h_vol *vol, uint32_t, double, double *, uint64_t) {
    if (!ctx || !vol) return fail(S2H_ERR_ARG, "null argument");
    return fail(S2H_ERR_UNSUPPORTED, "FP32 FFT path not built yet");
}
}
