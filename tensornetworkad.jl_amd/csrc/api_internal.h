// The context object behind tnad_ctx* and the error funnel of every extern "C" entry.
#pragma once
#include <exception>

#include "tnad_internal.h"

struct tnad_ctx {
  tnad::Ctx c;
  tnad_stats stats{};
};

namespace tnad {

// Runs f on the context's device, waits for the stream, and maps exceptions to the
// TNAD_ERR_* return codes (message kept for tnad_last_error).
template <typename F>
int guarded(tnad_ctx* ctx, F&& f) {
  if (!ctx) return TNAD_ERR_ARG;
  try {
    TNAD_HIP_TRY(hipSetDevice(ctx->c.device));
    f();
    ctx->c.sync();
    ctx->c.last_error.clear();
    return TNAD_OK;
  } catch (const Error& e) {
    ctx->c.last_error = e.msg;
    (void)hipGetLastError();
    return e.code;
  } catch (const std::exception& e) {
    ctx->c.last_error = e.what();
    return TNAD_ERR_ARG;
  }
}

}  // namespace tnad
