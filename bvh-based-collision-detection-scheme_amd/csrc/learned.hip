// Learned restructure scheduler kernels (placeholder until the MLP kernels land).
#include "learned.hpp"

namespace bvhcd {
struct LearnedState {
    int dummy = 0;
};
LearnedState* learned_create(hipStream_t) { return new LearnedState(); }
void learned_destroy(LearnedState* L) { delete L; }
static int nyi(std::string* err) {
    if (err) *err = "learned scheduler not built yet";
    return BVHCD_E_INVALID;
}
int learned_set_weights(LearnedState*, const float*, hipStream_t, std::string* err) { return nyi(err); }
int learned_get_weights(LearnedState*, float*, hipStream_t, std::string* err) { return nyi(err); }
int learned_features(LearnedState*, DevState&, int32_t, int64_t, float*, int32_t*, int64_t, int64_t*, hipStream_t,
                     std::string* err) { return nyi(err); }
int learned_forward(LearnedState*, const float*, int64_t, float*, int, hipStream_t, std::string* err) { return nyi(err); }
int learned_train_step(LearnedState*, const float*, const float*, int64_t, float, float*, int, hipStream_t,
                       std::string* err) { return nyi(err); }
int learned_grad(LearnedState*, const float*, const float*, int64_t, float*, float*, int, hipStream_t,
                 std::string* err) { return nyi(err); }
bool learned_ready(const LearnedState*) { return false; }
int learned_select(LearnedState*, DevState&, int32_t, int64_t, hipStream_t, int64_t*, int64_t*, int64_t*,
                   std::string* err) { return nyi(err); }
}  // namespace bvhcd
