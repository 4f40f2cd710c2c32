// cucm_serve.cpp -- serving loop (placeholder until continuous batching lands).
#include "cucm_engine.h"

struct Server {};

void serve_shutdown(cucm_ctx *c) {
  delete c->server;
  c->server = nullptr;
}

extern "C" {
int cucm_serve_start(cucm_ctx *ctx, uint32_t) { return ctx ? CUCM_E_UNSUPPORTED : CUCM_E_PARAM; }
int cucm_serve_stop(cucm_ctx *ctx) { return ctx ? CUCM_OK : CUCM_E_PARAM; }
int cucm_submit(cucm_ctx *ctx, const cucm_job *, uint64_t *) {
  return ctx ? CUCM_E_UNSUPPORTED : CUCM_E_PARAM;
}
int cucm_poll(cucm_ctx *ctx, uint64_t, int, cucm_result *) {
  return ctx ? CUCM_E_UNSUPPORTED : CUCM_E_PARAM;
}
}
