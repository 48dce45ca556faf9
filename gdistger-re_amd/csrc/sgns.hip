// sgns.hip - skip-gram negative-sampling trainer (filled in below).
#include "dg_internal.h"

struct dg_sgns_state {
  int placeholder = 0;
};

void dg_sgns_free(dg_ctx* ctx) {
  delete ctx->sgns;
  ctx->sgns = nullptr;
}
