// lu_mixed.hip — mixed-precision factorization (FP32 / BF16 working precision).
#include "lu.hpp"

namespace lsx {
int mixed_alloc(lsx_lu* lu) { return fail(lu->ctx, LSX_EINVAL, "mixed precision not available yet"); }
void mixed_free(lsx_lu*) {}
int mixed_getrf(lsx_lu* lu) { return fail(lu->ctx, LSX_EINVAL, "mixed precision not available yet"); }
int mixed_get_factors(lsx_lu* lu, double*, int64_t) { return fail(lu->ctx, LSX_EINVAL, "mixed precision not available yet"); }
int mixed_getrs_dev(lsx_lu* lu, int64_t, const double*, int64_t, double*, int64_t) {
    return fail(lu->ctx, LSX_EINVAL, "mixed precision not available yet");
}
}  // namespace lsx
