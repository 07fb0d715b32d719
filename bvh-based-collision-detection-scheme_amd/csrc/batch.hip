// Batched scenes (placeholder until the batched pipeline lands).
#include "bvhcd_internal.hpp"

struct bvhcd_batch {
    std::string err = "batched pipeline not built yet";
};

extern "C" {
int bvhcd_batch_create(const bvhcd_config*, int32_t, const int64_t*, const int64_t*, bvhcd_batch** out) {
    if (out) *out = nullptr;
    return BVHCD_E_INVALID;
}
int bvhcd_batch_destroy(bvhcd_batch* b) { delete b; return BVHCD_OK; }
int bvhcd_batch_set_topology(bvhcd_batch*, const int32_t*, int) { return BVHCD_E_INVALID; }
int bvhcd_batch_update_vertices(bvhcd_batch*, const float*, int) { return BVHCD_E_INVALID; }
int bvhcd_batch_step(bvhcd_batch*, int, int) { return BVHCD_E_INVALID; }
int bvhcd_batch_get_pairs(bvhcd_batch*, int32_t, int32_t*, int64_t, int64_t*) { return BVHCD_E_INVALID; }
int bvhcd_batch_pair_counts(bvhcd_batch*, int64_t*) { return BVHCD_E_INVALID; }
const char* bvhcd_batch_last_error(const bvhcd_batch* b) { return b ? b->err.c_str() : "null batch"; }
int bvhcd_autotune(bvhcd_ctx*) { return BVHCD_E_INVALID; }
int bvhcd_autotune_report(char* buf, int64_t cap) {
    if (buf && cap > 0) buf[0] = 0;
    return BVHCD_E_INVALID;
}
}
