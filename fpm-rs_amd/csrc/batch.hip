// batch.hip — batched form (placeholder until the fused sweep kernel lands).
#include "session.h"

extern "C" {
int fpmhip_batch_create(fpmhip_context*, fpmhip_session* const*, uint64_t, fpmhip_batch**) {
  return FPMHIP_UNSUPPORTED;
}
int fpmhip_batch_destroy(fpmhip_batch*) { return FPMHIP_OK; }
int fpmhip_batch_run(fpmhip_batch*, uint64_t, double*) { return FPMHIP_UNSUPPORTED; }
}
