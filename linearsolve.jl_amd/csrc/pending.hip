// pending.hip — entry points whose implementation is not in this build yet; each returns
// LSX_EINVAL with a message so callers fail loudly.
#include "lu.hpp"

using namespace lsx;

extern "C" {
#define PENDING(ctx, what) return fail((ctx), LSX_EINVAL, "%s is not implemented in this build", what)
int lsx_serve_create(lsx_ctx* ctx, size_t, int32_t, lsx_serve**) { PENDING(ctx, "lsx_serve_create"); }
int lsx_serve_destroy(lsx_serve*) { return LSX_EINVAL; }
int lsx_serve_submit(lsx_serve*, const lsx_request*, double*, uint64_t*) { return LSX_EINVAL; }
int lsx_serve_poll(lsx_serve*, uint64_t, lsx_result*) { return LSX_EINVAL; }
int lsx_serve_drain(lsx_serve*) { return LSX_EINVAL; }
double lsx_serve_clock(lsx_serve*) { return 0.0; }
}
