// Learned restructure scheduler for gfx950: per-subtree features, an MLP 16 -> 64 -> 64 -> 1
// (ReLU) that predicts the traversal-cost saving of rebuilding a subtree instead of refitting
// it, and its training.
//
// The reference's own decision rule (front-quality flags, collision/bvtt_front/
// BvttFrontGenerateKernels.cu:34-49, decided in collision/lbvh/LBvhFixedDeformable.cu:210-267 of
// littlemine/BVH-based-Collision-Detection-Scheme, Source/Project/Mine/) stays the fallback
// when no weights are loaded. Either way the subtrees chosen are rebuilt by the same
// restructure path, so the pair set cannot change; only the amount of later traversal work does.
//
// Candidates ("treelets"): maximal subtrees with at most kTreeletLeaves leaves.
// Features (16 per treelet):
//   0  log(sum of internal-node surface areas now / at the last rebuild)   (SAH growth)
//   1  front entries (internal + external) per leaf of the subtree; an entry counts for the
//      subtree holding its node's leftmost leaf
//   2  mean over leaves of log(leaf box area now / at the last rebuild)    (box expansion)
//   3  depth of the subtree root / 32
//   4  log2(leaf count) / 12
//   5  min(1, frames since the subtree was last rebuilt / 12)
//   6-15 fraction of leaves whose external-front count c falls in bin floor(log2(1 + c)) (last bin 9+)
// Numerics: weights rounded to bf16 (nearest even), layer inputs rounded to bf16 before the two
// hidden products, fp32 accumulation (v_mfma_f32_16x16x16_bf16 for inference), fp32 biases
// and output layer. Training: per-sample fp32 forward/backward of the MSE loss through the
// same rounding (straight-through), block-reduced gradients, Adam on fp32 master weights.
#include <cstring>
#include <mutex>
#include <vector>

#include "kernels.hpp"
#include "learned.hpp"

namespace bvhcd {

constexpr int kIn = 16, kH = 64;
// Cost of the incremental front repair that follows a rebuild, in KEEP work per front entry:
// classification, KEEP of the kept entries and compaction each read the whole front (measured
// on the 4-layer cloth: ~32 ms against ~5-8 ms per KEEP frame of the same front).
constexpr float kDefaultRepairCost = 4.0f;
constexpr int kOffW1 = 0, kOffB1 = kH * kIn, kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kH * kH, kOffW3 = kOffB2 + kH,
              kOffB3 = kOffW3 + kH, kNParam = kOffB3 + 1;
static_assert(kNParam == BVHCD_MLP_NPARAM, "parameter packing");
constexpr int32_t kTreeletLeaves = 4096;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return (uint16_t)((u >> 16) | 0x40u);  // NaN
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }
__device__ inline float rbf(float f) { return bf2f(f2bf(f)); }

struct LearnedState {
    bool ready = false;
    float* params = nullptr;   // fp32 master weights
    uint16_t* wbf = nullptr;   // bf16 copy (all entries; biases used from params)
    float* grad = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    float* loss = nullptr;
    int64_t adam_t = 0;
    float* x = nullptr;        // io scratch
    float* y = nullptr;
    float* t = nullptr;
    int64_t rows_cap = 0;
    int32_t* roots = nullptr;  // candidate treelet roots
    uint32_t* flags = nullptr;
    uint32_t* pos = nullptr;
    uint32_t* total = nullptr;
    uint32_t* cand_cnt = nullptr; // front entries per candidate (by the node's leftmost leaf)
    int32_t* cand_of_leaf = nullptr; // candidate holding each leaf, -1 none
    int32_t* cand_of_int = nullptr;  // candidate holding each internal node's leftmost leaf, -1 none
    uint32_t* lcnt = nullptr;     // external-front entries per leaf
    bool by_query = true;         // the current fronts are ordered by query (else by node)
    int64_t counted_topo = -1;    // topology version the counts were made on
    int64_t counted_version = -1; // front version the per-leaf counts belong to
    int64_t node_cap = 0;
    float* feats = nullptr;
    float* score = nullptr;       // the network's score of each candidate (selection frames)
    float threshold = 0.0f;
    float repair_cost = kDefaultRepairCost;  // cost gate (0 = off), front entries' KEEP work per entry
    float* gate = nullptr;                   // device: predicted saving of the selection
};

static bool ensure(float*& p, int64_t& cap, int64_t need) {
    if (need <= cap && p) return true;
    if (p) (void)hipFree(p);
    p = nullptr;
    if (hipMalloc(&p, sizeof(float) * (size_t)need) != hipSuccess) return false;
    cap = need;
    return true;
}

LearnedState* learned_create(hipStream_t) {
    LearnedState* L = new LearnedState();
    bool ok = hipMalloc(&L->params, sizeof(float) * kNParam) == hipSuccess &&
              hipMalloc(&L->wbf, sizeof(uint16_t) * kNParam) == hipSuccess &&
              hipMalloc(&L->grad, sizeof(float) * kNParam) == hipSuccess &&
              hipMalloc(&L->m, sizeof(float) * kNParam) == hipSuccess &&
              hipMalloc(&L->v, sizeof(float) * kNParam) == hipSuccess &&
              hipMalloc(&L->loss, sizeof(float) * 4) == hipSuccess;
    if (!ok) {
        learned_destroy(L);
        return nullptr;
    }
    return L;
}

void learned_destroy(LearnedState* L) {
    if (!L) return;
    for (void* p : {(void*)L->params, (void*)L->wbf, (void*)L->grad, (void*)L->m, (void*)L->v, (void*)L->loss,
                    (void*)L->x, (void*)L->y, (void*)L->t, (void*)L->roots, (void*)L->flags, (void*)L->pos,
                    (void*)L->total, (void*)L->feats, (void*)L->cand_cnt, (void*)L->cand_of_leaf, (void*)L->lcnt,
                    (void*)L->cand_of_int, (void*)L->gate, (void*)L->score})
        if (p) (void)hipFree(p);
    delete L;
}

bool learned_ready(const LearnedState* L) { return L && L->ready; }

__global__ void k_cast_bf16(const float* __restrict__ p, uint16_t* __restrict__ w, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) w[i] = f2bf(p[i]);
}

static int refresh_bf16(LearnedState* L, hipStream_t s) {
    hipLaunchKernelGGL(k_cast_bf16, dim3(grid_for(kNParam)), dim3(kBlock), 0, s, L->params, L->wbf, kNParam);
    return hipGetLastError() == hipSuccess ? BVHCD_OK : BVHCD_E_DEVICE;
}

int learned_set_weights(LearnedState* L, const float* p, hipStream_t s, std::string* err) {
    if (hipMemcpyAsync(L->params, p, sizeof(float) * kNParam, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemsetAsync(L->m, 0, sizeof(float) * kNParam, s) != hipSuccess ||
        hipMemsetAsync(L->v, 0, sizeof(float) * kNParam, s) != hipSuccess || refresh_bf16(L, s) != BVHCD_OK ||
        hipStreamSynchronize(s) != hipSuccess) {
        if (err) *err = "mlp_set_weights: device error";
        return BVHCD_E_DEVICE;
    }
    L->adam_t = 0;
    L->ready = true;
    return BVHCD_OK;
}

int learned_get_weights(LearnedState* L, float* p, hipStream_t s, std::string* err) {
    if (hipMemcpyAsync(p, L->params, sizeof(float) * kNParam, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        if (err) *err = "mlp_get_weights: device error";
        return BVHCD_E_DEVICE;
    }
    return BVHCD_OK;
}

// ---------------------------------------------------------------- inference (MFMA)
// One wave64 computes 16 rows. v_mfma_f32_16x16x16_bf16 operand layout (lane l, i = 0..3):
//   A[m][k]: m = l % 16, k = 4 (l / 16) + i;   B[k][n]: n = l % 16, k = 4 (l / 16) + i;
//   D[m][n]: n = l % 16, m = 4 (l / 16) + i.
__global__ void __launch_bounds__(kBlock) k_mlp_forward(const float* __restrict__ x, int64_t rows,
                                                        const uint16_t* __restrict__ wbf, const float* __restrict__ p,
                                                        float* __restrict__ y) {
    __shared__ float h1[kBlock / 64][16][kH + 1];
    __shared__ float h2[kBlock / 64][16][kH + 1];
    const int w = threadIdx.x >> 6, lane = lane_id();
    const int64_t row0 = ((int64_t)blockIdx.x * (kBlock / 64) + w) * 16;
    const int r = lane & 15, kq = lane >> 4;
    // layer 1: X[16 x 16] . W1^T[16 x 64]
    bf16x4 a;
    for (int i = 0; i < 4; ++i) {
        const int64_t row = row0 + r;
        a[i] = (short)(row < rows ? f2bf(x[row * kIn + 4 * kq + i]) : 0);
    }
    for (int t = 0; t < 4; ++t) {
        const int n = 16 * t + r;
        bf16x4 b;
        for (int i = 0; i < 4; ++i) b[i] = (short)wbf[kOffW1 + n * kIn + 4 * kq + i];
        f32x4 c = {0.f, 0.f, 0.f, 0.f};
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
        for (int i = 0; i < 4; ++i) h1[w][4 * kq + i][n] = fmaxf(c[i] + p[kOffB1 + n], 0.f);
    }
    __syncthreads();
    // layer 2: H1[16 x 64] (bf16) . W2^T[64 x 64]
    for (int t = 0; t < 4; ++t) {
        const int n = 16 * t + r;
        f32x4 c = {0.f, 0.f, 0.f, 0.f};
        for (int ks = 0; ks < 4; ++ks) {
            bf16x4 aa, b;
            for (int i = 0; i < 4; ++i) {
                aa[i] = (short)f2bf(h1[w][r][16 * ks + 4 * kq + i]);
                b[i] = (short)wbf[kOffW2 + n * kH + 16 * ks + 4 * kq + i];
            }
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(aa, b, c, 0, 0, 0);
        }
        for (int i = 0; i < 4; ++i) h2[w][4 * kq + i][n] = fmaxf(c[i] + p[kOffB2 + n], 0.f);
    }
    __syncthreads();
    // output layer: fp32 activations, bf16-rounded weights
    if (lane < 16) {
        const int64_t row = row0 + lane;
        if (row < rows) {
            float acc = 0.f;
            for (int j = 0; j < kH; ++j) acc += h2[w][lane][j] * bf2f(wbf[kOffW3 + j]);
            y[row] = acc + p[kOffB3];
        }
    }
}

static int stage_rows(LearnedState* L, int64_t rows) {
    int64_t cap = L->rows_cap;
    if (rows <= cap && L->x) return BVHCD_OK;
    int64_t c2 = std::max<int64_t>(rows, 4096);
    int64_t cx = 0, cy = 0, ct = 0;
    if (!ensure(L->x, cx, c2 * kIn) || !ensure(L->y, cy, c2) || !ensure(L->t, ct, c2)) return BVHCD_E_OOM;
    L->rows_cap = c2;
    return BVHCD_OK;
}

static void launch_forward(LearnedState* L, const float* x, int64_t rows, float* y, hipStream_t s) {
    const unsigned blocks = (unsigned)((rows + 63) / 64);
    if (rows > 0) hipLaunchKernelGGL(k_mlp_forward, dim3(blocks), dim3(kBlock), 0, s, x, rows, L->wbf, L->params, y);
}

int learned_forward(LearnedState* L, const float* feats, int64_t rows, float* out, int where, hipStream_t s,
                    std::string* err) {
    if (!L->ready) {
        if (err) *err = "no weights";
        return BVHCD_E_INVALID;
    }
    if (where == BVHCD_MEM_DEVICE) {
        launch_forward(L, feats, rows, out, s);
        return hipStreamSynchronize(s) == hipSuccess ? BVHCD_OK : BVHCD_E_DEVICE;
    }
    if (stage_rows(L, rows)) return BVHCD_E_OOM;
    if (hipMemcpyAsync(L->x, feats, sizeof(float) * kIn * rows, hipMemcpyHostToDevice, s) != hipSuccess)
        return BVHCD_E_DEVICE;
    launch_forward(L, L->x, rows, L->y, s);
    if (hipMemcpyAsync(out, L->y, sizeof(float) * rows, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        if (err) *err = "mlp_forward: device error";
        return BVHCD_E_DEVICE;
    }
    return BVHCD_OK;
}

// ---------------------------------------------------------------- allocation-free inference
// Staging for learned_forward_noalloc: kNoAllocRows feature rows in, then kNoAllocRows scores
// out. The array lives in the code object: each GPU holds its own copy from module load until
// the process ends, at the address hipGetSymbolAddress reports with that GPU current (host code
// never uses the array's name as a device address). Every call on a GPU shares that GPU's copy,
// so calls that overlap on one GPU race. The forward kernel writes every score it produces and
// the rows are copied in whole, so no part of the staging needs clearing between calls.
constexpr int64_t kNoAllocRows = BVHCD_MLP_NOALLOC_MAX_ROWS;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) float scratch[kNoAllocRows * (kIn + 1)];

static std::mutex g_scratch_mu;
static float* g_scratch_ptr[kMaxDevices] = {};

// Looks up the device address of `scratch` on `device` (made current by the caller), once per
// device.
int noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return BVHCD_E_DEVICE;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return BVHCD_OK;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return BVHCD_E_DEVICE;
    g_scratch_ptr[device] = static_cast<float*>(p);
    return BVHCD_OK;
}

int learned_forward_noalloc(LearnedState* L, int device, const float* feats, int64_t rows, float* out,
                            hipStream_t s, std::string* err) {
    if (!L || !L->ready) {
        if (err) *err = "no weights";
        return BVHCD_E_INVALID;
    }
    if (rows > kNoAllocRows) {
        if (err) *err = "mlp_forward_noalloc: more rows than BVHCD_MLP_NOALLOC_MAX_ROWS";
        return BVHCD_E_INVALID;
    }
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || (cur != device && hipSetDevice(device) != hipSuccess)) {
        if (err) *err = "mlp_forward_noalloc: cannot select the context's device";
        return BVHCD_E_DEVICE;
    }
    if (noalloc_scratch_init(device) != BVHCD_OK) {
        if (err) *err = "mlp_forward_noalloc: scratch address lookup failed";
        return BVHCD_E_DEVICE;
    }
    if (rows == 0) return BVHCD_OK;
    float* x = g_scratch_ptr[device];
    float* y = x + kNoAllocRows * kIn;
    if (hipMemcpyAsync(x, feats, sizeof(float) * kIn * rows, hipMemcpyHostToDevice, s) != hipSuccess)
        return BVHCD_E_DEVICE;
    launch_forward(L, x, rows, y, s);
    if (hipMemcpyAsync(out, y, sizeof(float) * rows, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        if (err) *err = "mlp_forward_noalloc: device error";
        return BVHCD_E_DEVICE;
    }
    return BVHCD_OK;
}

// ---------------------------------------------------------------- training (fp32)
constexpr int kTrainBlock = 64;

__global__ void __launch_bounds__(kTrainBlock) k_mlp_grad(const float* __restrict__ x, const float* __restrict__ tgt,
                                                          int64_t rows, const float* __restrict__ p,
                                                          float* __restrict__ grad, float* __restrict__ loss) {
    __shared__ float W1[kH * kIn], W2[kH * kH], W3[kH], B1[kH], B2[kH];
    __shared__ float g[kNParam];
    __shared__ float lsum;
    for (int i = threadIdx.x; i < kH * kIn; i += kTrainBlock) W1[i] = rbf(p[kOffW1 + i]);
    for (int i = threadIdx.x; i < kH * kH; i += kTrainBlock) W2[i] = rbf(p[kOffW2 + i]);
    for (int i = threadIdx.x; i < kH; i += kTrainBlock) {
        W3[i] = rbf(p[kOffW3 + i]);
        B1[i] = p[kOffB1 + i];
        B2[i] = p[kOffB2 + i];
    }
    for (int i = threadIdx.x; i < kNParam; i += kTrainBlock) g[i] = 0.f;
    if (threadIdx.x == 0) lsum = 0.f;
    __syncthreads();
    const float b3 = p[kOffB3];
    const float inv_n = 1.0f / (float)rows;
    for (int64_t row = (int64_t)blockIdx.x * kTrainBlock + threadIdx.x; row < rows;
         row += (int64_t)gridDim.x * kTrainBlock) {
        float xq[kIn];
        for (int i = 0; i < kIn; ++i) xq[i] = rbf(x[row * kIn + i]);
        float z1[kH], z2[kH];
        for (int j = 0; j < kH; ++j) {
            float acc = 0.f;
            for (int i = 0; i < kIn; ++i) acc += W1[j * kIn + i] * xq[i];
            z1[j] = acc + B1[j];
        }
        for (int j = 0; j < kH; ++j) {
            float acc = 0.f;
            for (int k = 0; k < kH; ++k) acc += W2[j * kH + k] * rbf(fmaxf(z1[k], 0.f));
            z2[j] = acc + B2[j];
        }
        float yv = b3;
        for (int j = 0; j < kH; ++j) yv += W3[j] * fmaxf(z2[j], 0.f);
        const float res = yv - tgt[row];
        atomicAdd(&lsum, res * res * inv_n);
        const float dy = 2.f * res * inv_n;
        atomicAdd(&g[kOffB3], dy);
        for (int j = 0; j < kH; ++j) {
            const float h2 = fmaxf(z2[j], 0.f);
            atomicAdd(&g[kOffW3 + j], dy * h2);
            z2[j] = z2[j] > 0.f ? dy * W3[j] : 0.f;  // now dz2
            atomicAdd(&g[kOffB2 + j], z2[j]);
        }
        for (int k = 0; k < kH; ++k) {
            const float h1q = rbf(fmaxf(z1[k], 0.f));
            float dh1 = 0.f;
            for (int j = 0; j < kH; ++j) {
                atomicAdd(&g[kOffW2 + j * kH + k], z2[j] * h1q);
                dh1 += W2[j * kH + k] * z2[j];
            }
            z1[k] = z1[k] > 0.f ? dh1 : 0.f;  // now dz1
            atomicAdd(&g[kOffB1 + k], z1[k]);
            for (int i = 0; i < kIn; ++i) atomicAdd(&g[kOffW1 + k * kIn + i], z1[k] * xq[i]);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kNParam; i += kTrainBlock)
        if (g[i] != 0.f) atomicAdd(&grad[i], g[i]);
    if (threadIdx.x == 0) atomicAdd(loss, lsum);
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       float lr, float bc1, float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kNParam) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + 1e-8f);
}

static int run_grad(LearnedState* L, const float* feats, const float* targets, int64_t rows, int where,
                    hipStream_t s) {
    const float* dx = feats;
    const float* dt = targets;
    if (where != BVHCD_MEM_DEVICE) {
        if (stage_rows(L, rows)) return BVHCD_E_OOM;
        if (hipMemcpyAsync(L->x, feats, sizeof(float) * kIn * rows, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(L->t, targets, sizeof(float) * rows, hipMemcpyHostToDevice, s) != hipSuccess)
            return BVHCD_E_DEVICE;
        dx = L->x;
        dt = L->t;
    }
    (void)hipMemsetAsync(L->grad, 0, sizeof(float) * kNParam, s);
    (void)hipMemsetAsync(L->loss, 0, sizeof(float), s);
    unsigned blocks = (unsigned)((rows + kTrainBlock - 1) / kTrainBlock);
    if (blocks > 512) blocks = 512;
    hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(kTrainBlock), 0, s, dx, dt, rows, L->params, L->grad, L->loss);
    return hipGetLastError() == hipSuccess ? BVHCD_OK : BVHCD_E_DEVICE;
}

int learned_grad(LearnedState* L, const float* feats, const float* targets, int64_t rows, float* grad, float* loss,
                 int where, hipStream_t s, std::string* err) {
    int st = run_grad(L, feats, targets, rows, where, s);
    if (st) return st;
    float lv = 0;
    if (hipMemcpyAsync(grad, L->grad, sizeof(float) * kNParam, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(&lv, L->loss, sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        if (err) *err = "mlp_grad: device error";
        return BVHCD_E_DEVICE;
    }
    if (loss) *loss = lv;
    return BVHCD_OK;
}

int learned_train_step(LearnedState* L, const float* feats, const float* targets, int64_t rows, float lr, float* loss,
                       int where, hipStream_t s, std::string* err) {
    int st = run_grad(L, feats, targets, rows, where, s);
    if (st) return st;
    L->adam_t++;
    const float bc1 = 1.0f - powf(0.9f, (float)L->adam_t), bc2 = 1.0f - powf(0.999f, (float)L->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(grid_for(kNParam)), dim3(kBlock), 0, s, L->params, L->grad, L->m, L->v, lr, bc1,
                       bc2);
    if (refresh_bf16(L, s)) return BVHCD_E_DEVICE;
    float lv = 0;
    if (hipMemcpyAsync(&lv, L->loss, sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        if (err) *err = "mlp_train_step: device error";
        return BVHCD_E_DEVICE;
    }
    if (loss) *loss = lv;
    L->ready = true;
    return BVHCD_OK;
}

// ---------------------------------------------------------------- surface-area references
__device__ inline float rec_area(const NodeRec& r) {
    const float dx = fmaxf(r.hi.x - r.lo.x, 0.f), dy = fmaxf(r.hi.y - r.lo.y, 0.f), dz = fmaxf(r.hi.z - r.lo.z, 0.f);
    return 2.f * (dx * dy + dy * dz + dz * dx);
}

// all = 1: every node (after a build); else only nodes inside restructured runs.
__global__ void k_sa_ref(int32_t n, int all, int32_t frame, DevState d) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool leaf_in = all || d.rt_of_leaf[i] >= 0;
    if (leaf_in) {
        d.sa_ref_leaf[i] = rec_area(d.leaf[i]);
        d.built_frame[i] = frame;
    }
    if (i < n - 1) {
        bool in = all != 0;
        if (!in) {
            const int32_t r = d.rt_of_leaf[d.int_rx[i]];
            in = r >= 0 && i >= r && i < r + (d.int_ry[r] - d.int_rx[r]);
        }
        if (in) d.sa_ref[i] = rec_area(d.inode[i]);
    }
}

void update_sa_ref(DevState& d, int32_t n, int all, int32_t frame, hipStream_t s) {
    hipLaunchKernelGGL(k_sa_ref, dim3(grid_for(n)), dim3(kBlock), 0, s, n, all, frame, d);
}

// ---------------------------------------------------------------- candidates and features
__global__ void k_treelet_flags(int32_t I, DevState d, uint32_t* __restrict__ flags) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= I) return;
    const int32_t cnt = d.int_ry[i] - d.int_rx[i] + 1;
    const int32_t p = d.int_par[i];
    const bool root = cnt <= kTreeletLeaves && (p < 0 || d.int_ry[p] - d.int_rx[p] + 1 > kTreeletLeaves);
    flags[i] = root ? 1u : 0u;
}
__global__ void k_treelet_compact(int32_t I, const uint32_t* __restrict__ flags, const uint32_t* __restrict__ pos,
                                  int32_t* __restrict__ roots) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < I && flags[i]) roots[pos[i]] = i;
}

// Candidate of every leaf: -1, then candidate c (roots in preorder, compacted) over its leaves.
__global__ void __launch_bounds__(kBlock) k_cand_of_leaf(const int32_t* __restrict__ roots, DevState d,
                                                         int32_t* __restrict__ cand_of_leaf) {
    const int32_t r = roots[blockIdx.x];
    for (int32_t j = d.int_rx[r] + threadIdx.x; j <= d.int_ry[r]; j += kBlock) cand_of_leaf[j] = blockIdx.x;
}
// The candidate of every internal node's leftmost leaf, so that counting an internal front entry
// takes one lookup.
__global__ void __launch_bounds__(kBlock) k_cand_of_int(int32_t I, DevState d, const int32_t* __restrict__ cand_of_leaf,
                                                        int32_t* __restrict__ cand_of_int) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < I) cand_of_int[i] = cand_of_leaf[d.int_rx[i]];
}

// Front counts for the features, recomputed only when the front changed:
//  * cand_cnt[c]: entries whose key leaf (an internal entry's node's leftmost leaf, an external
//    entry's leaf) lies in candidate c;
//  * lcnt[j]: external entries at leaf j.
// Each workgroup takes one contiguous piece of the fronts and counts in LDS: candidates in a
// table of all of them (when they fit), leaves in a window starting at the piece's first query
// (query-ordered fronts: a query's external leaves follow it closely in Morton order) or first
// leaf (node-ordered fronts); keys outside go to global atomics. The tables are then added to
// the global counts once per non-zero bin.
constexpr int kCandLds = 4096, kLeafWin = 4096;
// Workgroups of the count: smaller pieces keep more of a piece's external leaves inside its
// LDS window (the rest take global atomics).
constexpr int kCountGrid = 2048;

// Adds one to ctr[key] for every lane where `valid` holds, with one atomic per run of equal keys
// in neighbouring lanes: on query-ordered fronts a wave's entries mostly share their candidate,
// and one atomic per lane on one word would serialise. Called by the whole wave.
template <class T>
__device__ inline void run_count(T* ctr, bool valid, uint32_t key) {
    const int lane = lane_id();
    const uint64_t vmask = __ballot(valid);
    const uint32_t prev = __shfl_up(key, 1);
    const bool prev_valid = lane > 0 && ((vmask >> (lane - 1)) & 1ull) != 0;
    const bool start = valid && !(prev_valid && prev == key);
    const uint64_t starts = __ballot(start);
    if (start) {  // the run ends at the next start or the next invalid lane
        const uint64_t le = (2ull << lane) - 1ull;
        const uint64_t stop = (starts | ~vmask) & ~le;
        const int e = stop ? __ffsll((long long)stop) - 1 : 64;
        atomicAdd(ctr + key, (uint32_t)(e - lane));
    }
}
__global__ void __launch_bounds__(kBlock) k_count_front(DevState d, const int32_t* __restrict__ cand_of_leaf,
                                                        const int32_t* __restrict__ cand_of_int,
                                                        int32_t ncand, int by_query, uint32_t* __restrict__ cand_cnt,
                                                        uint32_t* __restrict__ lcnt) {
    __shared__ uint32_t hc[kCandLds];
    __shared__ uint32_t hl[kLeafWin];
    __shared__ int32_t wbase;
    const bool lds_cand = ncand <= kCandLds;
    for (int i = threadIdx.x; i < kCandLds; i += kBlock) hc[i] = 0u;
    for (int i = threadIdx.x; i < kLeafWin; i += kBlock) hl[i] = 0u;
    const int c0 = d.fcur[0], c1 = d.fcur[1];
    const uint32_t ni = (uint32_t)min((int64_t)d.fsize[c0], d.front_cap[0]);
    const uint32_t ne = (uint32_t)min((int64_t)d.fsize[2 + c1], d.front_cap[1]);
    const uint32_t total = ni + ne;
    const uint32_t piece = (total + gridDim.x - 1) / gridDim.x;
    const uint32_t k0 = min(blockIdx.x * piece, total), k1 = min(k0 + piece, total);
    if (threadIdx.x == 0) {
        const uint32_t e0 = max(k0, ni);  // first external entry of the piece
        wbase = e0 < k1 ? (by_query ? d.fquery[1][c1][e0 - ni] + 1 : (int32_t)d.fnode[1][c1][e0 - ni]) : 0;
    }
    __syncthreads();
    const int32_t wb = wbase;
    // kCountU entries per thread and iteration (an entry, then its candidate: an external entry's
    // leaf's, an internal entry's node's), their loads issued together
    constexpr int kCountU = 4;
    for (uint32_t kb = k0; kb < k1; kb += kCountU * kBlock) {  // wave-uniform trip count
        int32_t nd[kCountU];  // the entry's node: an external entry's leaf
#pragma unroll
        for (int u = 0; u < kCountU; ++u) {
            const uint32_t k = kb + u * kBlock + threadIdx.x;
            nd[u] = k < k1 ? (int32_t)(k >= ni ? d.fnode[1][c1][k - ni] : d.fnode[0][c0][k]) : 0;
        }
        int32_t c[kCountU];
#pragma unroll
        for (int u = 0; u < kCountU; ++u) {
            const uint32_t k = kb + u * kBlock + threadIdx.x;
            c[u] = k < k1 ? (k >= ni ? cand_of_leaf[nd[u]] : cand_of_int[nd[u]]) : -1;
        }
#pragma unroll
        for (int u = 0; u < kCountU; ++u) {
            const uint32_t k = kb + u * kBlock + threadIdx.x;
            const bool in = k < k1, ext = k >= ni;
            if (lds_cand) run_count(hc, c[u] >= 0, (uint32_t)max(c[u], 0));
            else run_count(cand_cnt, c[u] >= 0, (uint32_t)max(c[u], 0));
            const uint32_t o = (uint32_t)(nd[u] - wb);
            const bool win = in && ext && o < (uint32_t)kLeafWin;
            run_count(hl, win, win ? o : 0u);
            if (in && ext && !win) atomicAdd(&lcnt[nd[u]], 1u);
        }
    }
    __syncthreads();
    if (lds_cand)
        for (int i = threadIdx.x; i < ncand; i += kBlock)
            if (hc[i]) atomicAdd(&cand_cnt[i], hc[i]);
    for (int i = threadIdx.x; i < kLeafWin; i += kBlock)
        if (hl[i]) atomicAdd(&lcnt[wb + i], hl[i]);
}

// One workgroup per treelet.
__global__ void __launch_bounds__(kBlock) k_features(const int32_t* __restrict__ roots, DevState d, int32_t frame,
                                                     float frame_scale, const uint32_t* __restrict__ cand_cnt,
                                                     const uint32_t* __restrict__ lcnt, float* __restrict__ feats) {
    __shared__ float red[4][kBlock];
    __shared__ uint32_t hist[10];
    const int32_t r = roots[blockIdx.x];
    const int32_t a = d.int_rx[r], b = d.int_ry[r], m = b - a + 1;
    if (threadIdx.x < 10) hist[threadIdx.x] = 0;
    __syncthreads();
    float sa_now = 0.f, sa_ref = 0.f, leaf_exp = 0.f, age = 0.f;
    for (int32_t i = r + threadIdx.x; i < r + m - 1; i += kBlock) {
        sa_now += rec_area(d.inode[i]);
        sa_ref += d.sa_ref[i];
    }
    for (int32_t j = a + threadIdx.x; j <= b; j += kBlock) {
        const float now = rec_area(d.leaf[j]), ref = d.sa_ref_leaf[j];
        leaf_exp += logf((now + 1e-12f) / (ref + 1e-12f));
        age += (float)(frame - d.built_frame[j]);
        const uint32_t c = lcnt[j];
        const int bin = c >= 511u ? 9 : 31 - __clz(c + 1u);  // floor(log2(1 + c)), c + 1 in [1, 511]
        atomicAdd(&hist[bin], 1u);
    }
    red[0][threadIdx.x] = sa_now;
    red[1][threadIdx.x] = sa_ref;
    red[2][threadIdx.x] = leaf_exp;
    red[3][threadIdx.x] = age;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        float* f = feats + (int64_t)blockIdx.x * kIn;
        f[0] = logf((red[0][0] + 1e-12f) / (red[1][0] + 1e-12f));
        f[1] = (float)cand_cnt[blockIdx.x] / (float)m;
        f[2] = red[2][0] / (float)m;
        int depth = 0;
        for (int32_t p = d.int_par[r]; p >= 0 && depth < 4096; p = d.int_par[p]) ++depth;
        f[3] = (float)depth / 32.f;
        f[4] = log2f((float)m) / 12.f;
        f[5] = fminf(red[3][0] / (float)m * frame_scale, 1.0f);
        for (int k = 0; k < 10; ++k) f[6 + k] = (float)hist[k] / (float)m;
    }
}

int learned_reserve(LearnedState* L, int32_t n) {
    if (!L) return BVHCD_E_INVALID;
    if (L->node_cap >= n && L->gate) return BVHCD_OK;
    if (L->node_cap < n) {
        for (void* p : {(void*)L->roots, (void*)L->flags, (void*)L->pos, (void*)L->total, (void*)L->feats,
                        (void*)L->cand_cnt, (void*)L->cand_of_leaf, (void*)L->cand_of_int, (void*)L->lcnt,
                        (void*)L->score})
            if (p) (void)hipFree(p);
        L->roots = nullptr; L->flags = nullptr; L->pos = nullptr; L->total = nullptr; L->feats = nullptr;
        L->cand_cnt = nullptr; L->cand_of_leaf = nullptr; L->cand_of_int = nullptr; L->lcnt = nullptr;
        L->score = nullptr;
        L->node_cap = 0;
        L->counted_version = -1;
        // candidates are disjoint internal subtrees, so there are fewer than n of them
        if (hipMalloc(&L->cand_cnt, sizeof(uint32_t) * n) != hipSuccess ||
            hipMalloc(&L->cand_of_leaf, sizeof(int32_t) * n) != hipSuccess ||
            hipMalloc(&L->cand_of_int, sizeof(int32_t) * n) != hipSuccess ||
            hipMalloc(&L->lcnt, sizeof(uint32_t) * n) != hipSuccess ||
            hipMalloc(&L->roots, sizeof(int32_t) * n) != hipSuccess ||
            hipMalloc(&L->flags, sizeof(uint32_t) * n) != hipSuccess ||
            hipMalloc(&L->pos, sizeof(uint32_t) * n) != hipSuccess ||
            hipMalloc(&L->total, sizeof(uint32_t) * 2) != hipSuccess ||
            hipMalloc(&L->feats, sizeof(float) * kIn * (size_t)n) != hipSuccess ||
            hipMalloc(&L->score, sizeof(float) * (size_t)n) != hipSuccess)
            return BVHCD_E_OOM;
        L->node_cap = n;
    }
    if (!L->gate && hipMalloc(&L->gate, sizeof(float) * 2) != hipSuccess) return BVHCD_E_OOM;
    return BVHCD_OK;
}

// candidate roots + features on the device; returns the row count (synchronises)
static int compute_features(LearnedState* L, DevState& d, int32_t n, int64_t frame, hipStream_t s, int64_t* rows) {
    const int32_t I = n - 1;
    *rows = 0;
    if (I < 1) return BVHCD_OK;
    if (int st = learned_reserve(L, n)) return st;
    hipLaunchKernelGGL(k_treelet_flags, dim3(grid_for(I)), dim3(kBlock), 0, s, I, d, L->flags);
    launch_exscan(d, L->flags, L->pos, I, L->total, s);
    hipLaunchKernelGGL(k_treelet_compact, dim3(grid_for(I)), dim3(kBlock), 0, s, I, L->flags, L->pos, L->roots);
    uint32_t cnt = 0;
    if (hipMemcpyAsync(&cnt, L->total, sizeof(cnt), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return BVHCD_E_DEVICE;
    // front counts of this front on these candidates (the candidates change with the topology)
    if (cnt > 0 && (L->counted_version != d.front_version || L->counted_topo != d.topo_version)) {
        L->by_query = d.front_by_query != 0;
        fill_i32(L->cand_of_leaf, n, -1, s);
        hipLaunchKernelGGL(k_cand_of_leaf, dim3(cnt), dim3(kBlock), 0, s, L->roots, d, L->cand_of_leaf);
        hipLaunchKernelGGL(k_cand_of_int, dim3(grid_for(I)), dim3(kBlock), 0, s, I, d, L->cand_of_leaf,
                           L->cand_of_int);
        hipMemsetAsync(L->cand_cnt, 0, sizeof(uint32_t) * (size_t)cnt, s);
        hipMemsetAsync(L->lcnt, 0, sizeof(uint32_t) * (size_t)n, s);
        hipLaunchKernelGGL(k_count_front, dim3(kCountGrid), dim3(kBlock), 0, s, d, L->cand_of_leaf, L->cand_of_int,
                           (int32_t)cnt,
                           L->by_query ? 1 : 0, L->cand_cnt, L->lcnt);
        L->counted_version = d.front_version;
        L->counted_topo = d.topo_version;
    }
    if (cnt > 0)
        hipLaunchKernelGGL(k_features, dim3(cnt), dim3(kBlock), 0, s, L->roots, d, (int32_t)frame, 1.0f / 12.0f,
                           L->cand_cnt, L->lcnt, L->feats);
    *rows = cnt;
    return hipGetLastError() == hipSuccess ? BVHCD_OK : BVHCD_E_DEVICE;
}

int learned_features(LearnedState* L, DevState& d, int32_t n, int64_t frame, float* feats, int32_t* roots, int64_t cap,
                     int64_t* rows, hipStream_t s, std::string* err) {
    int st = compute_features(L, d, n, frame, s, rows);
    if (st) {
        if (err) *err = "mlp_features: device error";
        return st;
    }
    const int64_t k = std::min<int64_t>(cap, *rows);
    if (k > 0) {
        if ((feats && hipMemcpyAsync(feats, L->feats, sizeof(float) * kIn * k, hipMemcpyDeviceToHost, s) != hipSuccess) ||
            (roots && hipMemcpyAsync(roots, L->roots, sizeof(int32_t) * k, hipMemcpyDeviceToHost, s) != hipSuccess))
            return BVHCD_E_DEVICE;
    }
    return hipStreamSynchronize(s) == hipSuccess ? BVHCD_OK : BVHCD_E_DEVICE;
}

// ---------------------------------------------------------------- selection
// Marks the chosen roots like the heuristic selection does (difference array of root + 1 over
// the leaf range, counters of subtrees / internal nodes / leaves).
__global__ void k_mark_roots(int64_t rows, const int32_t* __restrict__ roots, const float* __restrict__ pred,
                             float threshold, int use_pred, DevState d) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= rows) return;
    if (use_pred && !(pred[k] > threshold)) return;
    const int32_t r = roots[k];
    const int32_t a = d.int_rx[r], b = d.int_ry[r];
    atomicAdd(&d.ctr[C_RT_SUB], 1u);
    atomicAdd(&d.ctr[C_RT_INT], (uint32_t)(b - a));
    atomicAdd(&d.ctr[C_RT_LEAF], (uint32_t)(b - a + 1));
    atomicAdd(&d.r_leaf_lo[a], r + 1);
    atomicAdd(&d.r_leaf_lo[b + 1], -(r + 1));
}

static int read_counts(DevState& d, hipStream_t s, int64_t* nsub, int64_t* nint, int64_t* nleaf) {
    uint32_t cnt[3];
    if (hipMemcpyAsync(cnt, d.ctr + C_RT_SUB, sizeof(cnt), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return BVHCD_E_DEVICE;
    *nsub = cnt[0];
    *nint = cnt[1];
    *nleaf = cnt[2];
    return BVHCD_OK;
}

// Predicted saving of a selection, in front entries per frame: sum over the selected
// candidates of (predicted fraction of their entries saved) x (their entries = feature 1 x
// leaves). One workgroup.
__global__ void __launch_bounds__(kBlock) k_gate(int64_t rows, const int32_t* __restrict__ roots,
                                                 const float* __restrict__ pred, float threshold,
                                                 const float* __restrict__ feats, DevState d, float* out) {
    __shared__ float red[kBlock];
    float acc = 0.f;
    for (int64_t k = threadIdx.x; k < rows; k += kBlock) {
        const float y = pred[k];
        if (!(y > threshold)) continue;
        const int32_t r = roots[k];
        const float leaves = (float)(d.int_ry[r] - d.int_rx[r] + 1);
        acc += y * feats[k * kIn + 1] * leaves;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// The network ranks subtrees; the cost gate judges the whole selection: a rebuild is followed
// by an incremental front repair whose passes read the whole front (classification, KEEP of
// the kept entries, compaction), about repair_cost x (front entries) of KEEP work, so the
// selection is applied only if its predicted saving over `horizon` frames pays for that.
int learned_select(LearnedState* L, DevState& d, int32_t n, int64_t frame, int32_t horizon, hipStream_t s,
                   int64_t* nsub, int64_t* nint, int64_t* nleaf, bool* rejected, std::string* err) {
    *rejected = false;
    int64_t rows = 0;
    int st = compute_features(L, d, n, frame, s, &rows);
    if (st) {
        if (err) *err = "learned selection: features failed";
        return st;
    }
    if ((st = learned_reserve(L, n))) return st;  // allocated already unless n < 2 (no candidates)
    (void)hipMemsetAsync(d.ctr + C_RT_SUB, 0, sizeof(uint32_t) * 3, s);
    (void)hipMemsetAsync(d.r_leaf_lo, 0, sizeof(int32_t) * ((size_t)n + 1), s);
    (void)hipMemsetAsync(L->gate, 0, sizeof(float) * 2, s);
    if (rows > 0) {
        launch_forward(L, L->feats, rows, L->score, s);
        hipLaunchKernelGGL(k_mark_roots, dim3(grid_for(rows)), dim3(kBlock), 0, s, rows, L->roots, L->score,
                           L->threshold, 1, d);
        hipLaunchKernelGGL(k_gate, dim3(1), dim3(kBlock), 0, s, rows, L->roots, L->score, L->threshold, L->feats, d,
                           L->gate);
    }
    float saving = 0.f;
    uint32_t sz[4] = {0, 0, 0, 0};
    if (hipMemcpyAsync(&saving, L->gate, sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(sz, d.fsize, sizeof(sz), hipMemcpyDeviceToHost, s) != hipSuccess)
        return BVHCD_E_DEVICE;
    st = read_counts(d, s, nsub, nint, nleaf);
    if (st) return st;
    const double entries = (double)sz[d.fcur[0]] + (double)sz[2 + d.fcur[1]];
    if (*nsub > 0 && L->repair_cost > 0.f && (double)saving * horizon < (double)L->repair_cost * entries) {
        *nsub = *nint = *nleaf = 0;  // not worth its front repair: refit instead
        *rejected = true;
    }
    return BVHCD_OK;
}

// Explicit selection (host-given roots), validated: internal ids with disjoint leaf ranges.
int mark_given_roots(DevState& d, int32_t n, const int32_t* roots_host, int64_t k, hipStream_t s, int64_t* nsub,
                     int64_t* nint, int64_t* nleaf, std::string* err) {
    if (k < 0 || k > n) return BVHCD_E_INVALID;
    std::vector<int32_t> rx(k), ry(k);
    std::vector<int32_t> ids(roots_host, roots_host + k);
    for (int32_t r : ids)
        if (r < 0 || r >= n - 1) {
            if (err) *err = "restructure: root id out of range";
            return BVHCD_E_INVALID;
        }
    int32_t* droots = d.r_base;  // scratch (T entries), overwritten later by the restructure emission
    // read ranges to check disjointness
    std::vector<int32_t> arx(n - 1 > 0 ? n - 1 : 1), ary(n - 1 > 0 ? n - 1 : 1);
    if (hipMemcpyAsync(arx.data(), d.int_rx, sizeof(int32_t) * (n - 1), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(ary.data(), d.int_ry, sizeof(int32_t) * (n - 1), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return BVHCD_E_DEVICE;
    std::vector<std::pair<int32_t, int32_t>> rg;
    for (int32_t r : ids) rg.push_back({arx[r], ary[r]});
    std::sort(rg.begin(), rg.end());
    for (size_t i = 1; i < rg.size(); ++i)
        if (rg[i].first <= rg[i - 1].second) {
            if (err) *err = "restructure: subtrees overlap";
            return BVHCD_E_INVALID;
        }
    (void)hipMemsetAsync(d.ctr + C_RT_SUB, 0, sizeof(uint32_t) * 3, s);
    (void)hipMemsetAsync(d.r_leaf_lo, 0, sizeof(int32_t) * ((size_t)n + 1), s);
    if (k > 0) {
        (void)hipMemcpyAsync(droots, ids.data(), sizeof(int32_t) * k, hipMemcpyHostToDevice, s);
        hipLaunchKernelGGL(k_mark_roots, dim3(grid_for(k)), dim3(kBlock), 0, s, k, droots, nullptr, 0.f, 0, d);
    }
    return read_counts(d, s, nsub, nint, nleaf);
}

void learned_set_threshold(LearnedState* L, float t) {
    if (L) L->threshold = t;
}
void learned_set_repair_cost(LearnedState* L, float cost) {
    if (L) L->repair_cost = cost;
}

}  // namespace bvhcd
