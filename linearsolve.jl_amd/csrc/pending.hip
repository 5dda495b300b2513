// pending.hip — entry points whose implementation is not in this build yet; each returns
// LSX_EINVAL with a message so callers fail loudly.
#include "lu.hpp"

using namespace lsx;

extern "C" {
#define PENDING(ctx, what) return fail((ctx), LSX_EINVAL, "%s is not implemented in this build", what)
int lsx_lu_refine(lsx_lu* lu, int64_t, const double*, int64_t, double*, int64_t, double, int32_t, int32_t,
                  lsx_refine_stats*) { PENDING(lu ? lu->ctx : nullptr, "lsx_lu_refine"); }
int lsx_batch_create(lsx_ctx* ctx, int64_t, const int64_t*, int, lsx_batch**) { PENDING(ctx, "lsx_batch_create"); }
int lsx_batch_destroy(lsx_batch*) { return LSX_EINVAL; }
int lsx_batch_factor_solve(lsx_batch*, const double*, const double*, double*, int64_t*, int64_t*) { return LSX_EINVAL; }
int lsx_csr_create(lsx_ctx* ctx, int64_t, int64_t, const int32_t*, const int32_t*, const double*, lsx_csr**) {
    PENDING(ctx, "lsx_csr_create");
}
int lsx_csr_destroy(lsx_csr*) { return LSX_EINVAL; }
int lsx_csr_spmv(lsx_csr*, const double*, double*) { return LSX_EINVAL; }
int lsx_gmres_solve(lsx_csr*, const double*, double*, int32_t, int32_t, double, double, int64_t, const double*,
                    const double*, lsx_gmres_stats*) { return LSX_EINVAL; }
int lsx_serve_create(lsx_ctx* ctx, size_t, int32_t, lsx_serve**) { PENDING(ctx, "lsx_serve_create"); }
int lsx_serve_destroy(lsx_serve*) { return LSX_EINVAL; }
int lsx_serve_submit(lsx_serve*, const lsx_request*, double*, uint64_t*) { return LSX_EINVAL; }
int lsx_serve_poll(lsx_serve*, uint64_t, lsx_result*) { return LSX_EINVAL; }
int lsx_serve_drain(lsx_serve*) { return LSX_EINVAL; }
double lsx_serve_clock(lsx_serve*) { return 0.0; }
}
