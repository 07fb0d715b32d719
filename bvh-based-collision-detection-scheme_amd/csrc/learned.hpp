// Learned restructure scheduler (MLP 16 -> 64 -> 64 -> 1, bf16 weights, fp32 accumulation).
#pragma once
#include <string>
#include "bvhcd_internal.hpp"

namespace bvhcd {
struct LearnedState;
LearnedState* learned_create(hipStream_t s);
void learned_destroy(LearnedState* L);
int learned_set_weights(LearnedState* L, const float* p, hipStream_t s, std::string* err);
int learned_get_weights(LearnedState* L, float* p, hipStream_t s, std::string* err);
int learned_features(LearnedState* L, DevState& d, int32_t n, int64_t frame, float* feats, int32_t* roots, int64_t cap,
                     int64_t* rows, hipStream_t s, std::string* err);
int learned_forward(LearnedState* L, const float* feats, int64_t rows, float* out, int where, hipStream_t s,
                    std::string* err);
int learned_train_step(LearnedState* L, const float* feats, const float* targets, int64_t rows, float lr, float* loss,
                       int where, hipStream_t s, std::string* err);
int learned_grad(LearnedState* L, const float* feats, const float* targets, int64_t rows, float* grad, float* loss,
                 int where, hipStream_t s, std::string* err);
// True once weights are loaded.
bool learned_ready(const LearnedState* L);
// Allocates the per-node decision buffers for a tree of n leaves (no-op when they fit), so that
// decision frames allocate nothing.
int learned_reserve(LearnedState* L, int32_t n);
// Selects the subtrees to restructure with the MLP (difference marks in d.r_leaf_lo, as the
// heuristic selection leaves them); returns subtree / internal-node / leaf counts.
// Scores the candidate subtrees and marks those above the threshold; the selection is dropped
// (*rejected, counts 0) when its predicted saving over `horizon` frames does not pay for the
// front repair (cost gate, off with repair cost 0).
int learned_select(LearnedState* L, DevState& d, int32_t n, int64_t frame, int32_t horizon, hipStream_t s,
                   int64_t* nsub, int64_t* nint, int64_t* nleaf, bool* rejected, std::string* err);
// Surface-area references and rebuild frame (all nodes after a build; all = 0: only the
// subtrees just restructured, from d.rt_of_leaf).
void update_sa_ref(DevState& d, int32_t n, int all, int32_t frame, hipStream_t s);
// Marks host-given subtree roots for restructure_subtrees (validated: disjoint ranges).
int mark_given_roots(DevState& d, int32_t n, const int32_t* roots, int64_t k, hipStream_t s, int64_t* nsub,
                     int64_t* nint, int64_t* nleaf, std::string* err);
// Forward pass over host rows staged through the library's static device scratch (no device
// allocation); rows <= BVHCD_MLP_NOALLOC_MAX_ROWS. noalloc_scratch_init looks the scratch up on
// a device (current), once per device.
int noalloc_scratch_init(int device);
int learned_forward_noalloc(LearnedState* L, int device, const float* feats, int64_t rows, float* out,
                            hipStream_t s, std::string* err);
void learned_set_threshold(LearnedState* L, float t);
void learned_set_repair_cost(LearnedState* L, float cost);
}  // namespace bvhcd
