// server.cpp — request loop (continuous batching over resident scenes). Filled in below the context lifetime.
#include "bz_ctx.h"

namespace bz {

struct Server {};

Ctx::Ctx() = default;
Ctx::~Ctx() = default;

}  // namespace bz

extern "C" {

bz_status bz_server_start(bz_ctx* ctx, const bz_server_cfg*) { return ctx ? ctx->fail(BZ_E_STATE, "server not available") : BZ_E_INVALID; }
bz_status bz_server_stop(bz_ctx* ctx) { return ctx ? BZ_OK : BZ_E_INVALID; }
bz_status bz_server_submit(bz_ctx* ctx, const bz_request*, uint64_t*) { return ctx ? BZ_E_STATE : BZ_E_INVALID; }
bz_status bz_server_poll(bz_ctx* ctx, uint64_t, bz_result*) { return ctx ? BZ_E_STATE : BZ_E_INVALID; }
bz_status bz_server_wait(bz_ctx* ctx, uint64_t, uint64_t, bz_result*) { return ctx ? BZ_E_STATE : BZ_E_INVALID; }
bz_status bz_server_fetch(bz_ctx* ctx, uint64_t, bz_IndirectDrawData*, uint32_t, uint32_t*) { return ctx ? BZ_E_STATE : BZ_E_INVALID; }
bz_status bz_server_cancel(bz_ctx* ctx, uint64_t) { return ctx ? BZ_E_STATE : BZ_E_INVALID; }
bz_status bz_server_release(bz_ctx* ctx, uint64_t) { return ctx ? BZ_E_STATE : BZ_E_INVALID; }
bz_status bz_server_get_stats(bz_ctx* ctx, bz_server_stats*) { return ctx ? BZ_E_STATE : BZ_E_INVALID; }

}
