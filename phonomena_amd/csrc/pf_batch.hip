// Batched stepping of many contexts (placeholder until batching lands).
#include "pf_runtime.hpp"

using namespace pf;

struct pf_batch { int precision; };

extern "C" int pf_batch_create(int32_t precision, pf_batch** out) {
    (void)precision; if (out) *out = nullptr;
    return fail(PF_ERR_STATE, "batching not available in this build");
}
extern "C" int pf_batch_add(pf_batch*, pf_ctx*) { return fail(PF_ERR_STATE, "batching not available"); }
extern "C" int pf_batch_remove(pf_batch*, pf_ctx*) { return fail(PF_ERR_STATE, "batching not available"); }
extern "C" int pf_batch_step(pf_batch*, int64_t) { return fail(PF_ERR_STATE, "batching not available"); }
extern "C" int pf_batch_sync(pf_batch*) { return fail(PF_ERR_STATE, "batching not available"); }
extern "C" int pf_batch_set_tiles(pf_batch*, int32_t, int32_t, int32_t) { return fail(PF_ERR_STATE, "batching not available"); }
extern "C" int pf_batch_destroy(pf_batch*) { return PF_OK; }
