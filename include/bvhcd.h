/*
 * bvhcd.h — C ABI of the MI355X (gfx950) BVH / BVTT-front collision-detection pipeline.
 *
 * Scene in (faces + per-frame vertices) -> LBVH -> BVTT front -> colliding triangle pairs out.
 * Plain pointers and sizes only; no C++ or HIP types cross this boundary.
 *
 * Reference interfaces each group replaces (reference paths are relative to
 * Source/Project/Mine/ of littlemine/BVH-based-Collision-Detection-Scheme):
 *   bvhcd_create / bvhcd_destroy
 *       LBvhFixedDeformable::LBvhFixedDeformable(LBvhBuildConfig&&)      collision/lbvh/LBvhFixedDeformable.h:30
 *       BvttFrontLooseIntra::BvttFrontLooseIntra(BvttFrontIntraBuildConfig) collision/bvtt_front/BvttFrontLooseIntra.h:31
 *       (capacities: setting/BvhSettings.h:49-54, setting/BvttFrontSettings.h:108-119)
 *   bvhcd_set_topology / bvhcd_update_vertices
 *       LBvhFixedDeformable::updatePrimData(const SceneData&)            collision/lbvh/LBvhFixedDeformable.cu:115-119
 *       (SceneData{pos, fids}: world/Scene.h:16-20)
 *   bvhcd_step(bvh_op, front_op)
 *       LBvhFixedDeformable::maintain(LBvhFixedDeformableMaintenance, SceneData)  collision/lbvh/LBvhFixedDeformable.cu:68-87
 *       BvttFrontLooseIntra::maintain(BvttFrontLooseIntraMaintenance)            collision/bvtt_front/BvttFrontLooseIntra.cu:69-126
 *   bvhcd_step_scheduled
 *       BenchmarkLogic::maintainScheme()                                  system/Logic/BenchmarkLogic.hpp:90-156
 *   bvhcd_get_pairs
 *       BvttFrontLooseIntra::applyCpResults(uint*, uint2*)                collision/bvtt_front/BvttFrontLooseIntra.cu:63-67
 *   bvhcd_get_stats
 *       BvttFrontLooseIntra::inspectResults()                             collision/bvtt_front/BvttFrontLooseIntra.cu:50-61
 *   bvhcd_export_bvh / bvhcd_export_front
 *       accessors LBvhFixedDeformable::{clvs,ctks,getPrimMap} and the front MultiArray
 *       (collision/lbvh/LBvhFixedDeformable.h:34-60, base/MultiArray.h:20-161) — exported so that
 *       callers and tests can check topology and front contents.
 *   bvhcd_scene_*
 *       world/Scene.cpp:14-49 (mesh ingest); here seeded synthetic deforming meshes.
 *
 * Every call returns 0 on success or a negative BVHCD_E* code; bvhcd_last_error() gives text.
 * Nothing exits or prints. Capacity overflow is detected on the device before any
 * out-of-bounds write; the step then fails with BVHCD_E_CAPACITY and bvhcd_get_stats()
 * reports the required sizes.
 */
#ifndef BVHCD_H_
#define BVHCD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define BVHCD_OK            0
#define BVHCD_E_INVALID    -1   /* bad argument or call order */
#define BVHCD_E_OOM        -2   /* device allocation failed */
#define BVHCD_E_CAPACITY   -3   /* front or pair capacity exceeded (see stats.required_*) */
#define BVHCD_E_DEVICE     -4   /* HIP runtime error */
#define BVHCD_E_NOT_BUILT  -5   /* step needs a BVH (or front) that does not exist yet */

/* ---- memory location of caller buffers ---- */
#define BVHCD_MEM_HOST   0
#define BVHCD_MEM_DEVICE 1
/* bvhcd_update_vertices only: the library reads the caller's device buffer in place (no copy)
   until the next bvhcd_update_vertices; the buffer must stay allocated and unchanged while the
   steps that read it run (they are ordered on the context stream). */
#define BVHCD_MEM_DEVICE_BORROW 2

/* ---- BVH maintenance op (LBvhFixedDeformableMaintenance, LBvhFixedDeformable.h:18-24) ---- */
#define BVHCD_BVH_BUILD       0
#define BVHCD_BVH_REFIT       1
#define BVHCD_BVH_RESTRUCTURE 2   /* restructure the subtrees flagged by the last quality check */
#define BVHCD_BVH_NONE        3   /* leave the BVH untouched (vertices must be unchanged) */

/* ---- front maintenance op (BvttFrontLooseIntraMaintenance, BvttFrontLooseIntra.h:19-26) ---- */
#define BVHCD_FRONT_PURE_BVH_CD 0   /* stackless self-CD traversal, no front kept */
#define BVHCD_FRONT_GENERATE    1
#define BVHCD_FRONT_UPDATE      2   /* prune + sprout, then quality check and ordering */
#define BVHCD_FRONT_REORDER     3   /* "balance": keep + re-bucket int front */
#define BVHCD_FRONT_KEEP        4
#define BVHCD_FRONT_RESTRUCTURE 5   /* incremental front restructure (BvttFrontLooseIntra::restructure,
                                       BvttFrontLooseIntra.cu:330-400): run by the library on the step
                                       after a repairing subtree rebuild; not accepted by bvhcd_step */

/* ---- scheduler (bvhcd_step_scheduled) ----
   The plain schemes are the reference's CDSchemeType cases of BenchmarkLogic::maintainScheme
   (system/Logic/BenchmarkLogic.hpp:90-156) and the ARCSim schemes of getSchemeOpt (:56-87),
   counted from the first frame after bvhcd_set_topology. In every plain scheme a REFIT frame
   after a flagging quality check becomes a RESTRUCTURE (LBvhFixedDeformable.cu:68-87) when
   enable_restructure is set; a BUILD frame always rebuilds. */
#define BVHCD_SCHED_PLAIN   0   /* STATIC_MANDATORY with divergent mark (the reference's active setting):
                                   BUILD every bvh_cycle frames; front GENERATE / UPDATE at front-cycle
                                   starts, REORDER at their ends, KEEP otherwise */
#define BVHCD_SCHED_LEARNED 1   /* MLP per-subtree restructure choice, heuristic fallback; the
                                   whole tree rebuilt as in BVHCD_SCHED_LEARNED_NO_NET */
#define BVHCD_SCHED_GENERATE 2  /* GENERATE scheme: BUILD + PURE_BVH_CD each frame */
#define BVHCD_SCHED_FRONT_GENERATE 3  /* FRONT_GENERATE scheme: BUILD + front GENERATE each frame */
#define BVHCD_SCHED_STATIC_NODIV 4    /* STATIC_MANDATORY without divergent mark (no REORDER) */
#define BVHCD_SCHED_REFIT_ONLY_FRONT 5 /* REFIT_ONLY_FRONT with divergent mark: one BUILD, then REFIT;
                                   front cycle as in PLAIN (BenchmarkLogic.hpp:129-141, whose missing
                                   break falls into unreachable code; this is that branch's schedule) */
#define BVHCD_SCHED_REFIT_ONLY_FRONT_NODIV 6 /* REFIT_ONLY_FRONT without divergent mark (no REORDER) */
#define BVHCD_SCHED_ARCSIM_STATIC 7   /* getSchemeOpt scheme 0: one BUILD; front cycle with REORDER */
#define BVHCD_SCHED_ARCSIM_PERIODIC 8 /* getSchemeOpt scheme 1: BUILD every bvh_cycle frames, front
                                   GENERATE every front_cycle frames, KEEP otherwise */
#define BVHCD_SCHED_LEARNED_NO_NET 9  /* the learned schedule with the network off (no reference
                                   counterpart: the control that isolates what the MLP adds): REFIT +
                                   KEEP, and BUILD + GENERATE once the time KEEP frames have lost to
                                   the aging tree since the last rebuild reaches the measured time of
                                   a rebuild frame; no subtree is ever restructured */

typedef struct bvhcd_config {
    int64_t max_tris;        /* capacity in triangles (>= 1) */
    int64_t max_verts;       /* capacity in vertices */
    int64_t int_front_cap;   /* internal-front entries; 0 = derive from free HBM */
    int64_t ext_front_cap;   /* external-front entries; 0 = derive from free HBM */
    int64_t pair_cap;        /* broad-phase pair capacity; 0 = derive from free HBM */
    int32_t narrow_phase;    /* 1 = run the triangle-triangle SAT (CDBenchmarkSettings includeNarrowPhase) */
    int32_t scheduler;       /* BVHCD_SCHED_* used by bvhcd_step_scheduled */
    int32_t enable_restructure; /* 1 = quality check + restructure (CDBenchmarkSettings enableRestr) */
    int32_t bvh_cycle;       /* mandatory rebuild cycle, 0 = 12 (BvhSettings.h:45) */
    int32_t front_cycle;     /* front cycle, 0 = 6 (BvttFrontSettings.h:105) */
    int32_t device;          /* HIP device ordinal */
    void*   stream;          /* hipStream_t to run on; NULL = the library creates one */
    int32_t skip_adjacent;   /* 1 = never report pairs of triangles that share a vertex (the simulator
                                build's covertex filter at every pair emission, e.g.
                                BvttFrontKeepKernels.cu:33-35,52-55); 0 = report them (default: the
                                triangle test reports them as intersecting) */
    int32_t auto_grow;       /* 1 = a step whose fronts or pair lists overflow grows them (to 1.25x what the
                                frame asked for, within free HBM) and redoes its front maintenance as a
                                GENERATE on the same tree, so it never reports BVHCD_E_CAPACITY while
                                memory lasts; costs one host synchronisation per step. 0 = report
                                the overflow (no counterpart in the reference, whose fixed-size
                                buffers overflow silently, BvttFrontLooseIntra.cu:56-58) */
} bvhcd_config;

typedef struct bvhcd_stats {
    int64_t frame;             /* frames stepped so far */
    int32_t last_bvh_op;       /* op actually executed on the last step */
    int32_t last_front_op;
    int64_t int_front;         /* current internal-front length */
    int64_t ext_front;         /* current external-front length */
    int64_t bp_pairs;          /* broad-phase pairs of the last step (cpNum) */
    int64_t np_pairs;          /* narrow-phase hits of the last step (actualCpNum) */
    int64_t required_int_front;/* sizes the last step needed (meaningful after BVHCD_E_CAPACITY) */
    int64_t required_ext_front;
    int64_t required_pairs;
    int64_t node_visits;       /* node tests made by traversal / front kernels in the last step */
    int64_t restructured_subtrees; /* subtrees rebuilt by the last step (restructure) */
    int64_t restructured_leaves;
    double  stage_ms[8];       /* last step: 0 build, 1 refit, 2 restructure, 3 front/broad,
                                  4 ordering + quality, 5 narrow, 6 scheduler/MLP, 7 total */
    double  stage_bytes[8];    /* algorithmic HBM bytes of those stages in the last step */
    int64_t implicit_pairs;    /* of bp_pairs (and of np_pairs with a narrow phase): KEEP's pairs of
                                  triangles sharing a vertex, counted from the front's head without
                                  a test (every coordinate finite: both phases report them anyway) */
} bvhcd_stats;

typedef struct bvhcd_ctx bvhcd_ctx;

int  bvhcd_create(const bvhcd_config* cfg, bvhcd_ctx** out);
int  bvhcd_destroy(bvhcd_ctx* ctx);
const char* bvhcd_last_error(const bvhcd_ctx* ctx);

/* faces: T x int32[3] vertex ids. Topology is fixed until the next call (it resets BVH and front). */
int  bvhcd_set_topology(bvhcd_ctx* ctx, const int32_t* faces, int64_t num_tris, int64_t num_verts, int where);
/* xyz: V x float[3]. Device memory: copied D2D on the context stream. Host memory: uploaded on
   the context's copy stream into one of two device buffers, overlapping the step still running
   on the GPU; returns when the copy is complete, so xyz may be reused at once. */
int  bvhcd_update_vertices(bvhcd_ctx* ctx, const float* xyz, int64_t num_verts, int where);
/* Box mode, the simulator-facing input of the reference (ARCSimSceneData, world/Scene.h:39-46;
   LBvhFixedDeformable::maintain, LBvhFixedDeformable.cu:89-112; calcMaxBVARCSim / calcMCsARCSim,
   LBvhBuildKernels.cu:10-33; refitExtNodeARCSim, LBvhUpdateKernels.cu:10-33): per-triangle boxes
   (num_tris x 6 floats: min x y z, max x y z, in triangle order) replace the vertices for the
   scene box, the Morton codes, builds, restructures and refits of the following steps, until the
   next bvhcd_update_vertices. Boxes with min > max on an axis overlap nothing on it. The narrow
   phase needs vertices: steps in box mode return BVHCD_E_INVALID when it is enabled. */
int  bvhcd_update_boxes(bvhcd_ctx* ctx, const float* boxes, int64_t num_tris, int where);

/* One frame with explicit ops (LBvh::maintain then BvttFront::maintain). Stream-ordered. */
int  bvhcd_step(bvhcd_ctx* ctx, int bvh_op, int front_op);
/* One frame with ops chosen by the configured scheduler. */
int  bvhcd_step_scheduled(bvhcd_ctx* ctx);
/* The ops a plain scheme schedules at `frame` before any restructure override (host only, no
   context): BenchmarkLogic::maintainScheme / getSchemeOpt (system/Logic/BenchmarkLogic.hpp:56-156).
   The learned schedulers start from the PLAIN table. cycles <= 0: 12 and 6. */
int  bvhcd_schedule_ops(int sched, int64_t frame, int bvh_cycle, int front_cycle, int* bvh_op, int* front_op);

/* Pairs of the last step as (tri_a, tri_b) int32 pairs with tri_a < tri_b, unordered list.
   filtered=1: narrow-phase hits; 0: broad-phase (AABB) pairs. Synchronises the stream. */
int  bvhcd_get_pairs(bvhcd_ctx* ctx, int32_t* out, int64_t cap_pairs, int64_t* num_pairs, int where, int filtered);
/* Order-independent 64-bit digest of the pair set of the last step (sum of a 64-bit mix of
   each pair, with the count), computed on the device. */
int  bvhcd_pair_digest(bvhcd_ctx* ctx, int filtered, uint64_t* digest, int64_t* num_pairs);
int  bvhcd_get_stats(bvhcd_ctx* ctx, bvhcd_stats* out);
/* Per-stage hipEvent timing: frames > 0 records up to that many frames (events plus a
   device-side counter log, no host syncs); 0 switches it off. stats.stage_ms then holds the
   last step's stage times, and bvhcd_timing_frames() returns every recorded frame. */
int  bvhcd_set_timing(bvhcd_ctx* ctx, int frames);
/* Recorded frames since bvhcd_set_timing: per frame 8 stage times (ms, layout of
   bvhcd_stats.stage_ms), 8 algorithmic byte counts (layout of stage_bytes) and the executed
   (bvh_op, front_op). Synchronises the stream. */
int  bvhcd_timing_frames(bvhcd_ctx* ctx, int64_t cap, double* ms, double* bytes, int32_t* ops, int64_t* n);
/* The counters behind those byte counts, per recorded frame (8 values each): node tests,
   broad-phase pairs, narrow-phase hits, internal and external front length after the step,
   overflow flags, implicit shared-vertex pairs, of those narrow-phase hits (the reference's
   Logger record of front sizes and pair counts, BvttFrontLooseIntra.cu:50-61). */
int  bvhcd_timing_counts(bvhcd_ctx* ctx, int64_t cap, int64_t* counts, int64_t* n);

/* The recorded frames as the reference's log file text (Logger::record, system/Log/Logger.hpp:68-94):
   an empty CPU-timer block, then per frame the operation messages and the inspectResults line
   "front (int, ext) cpNum bp actualCpNum np" (BvttFrontLooseIntra.cu:50-61), then per frame one
   "@section: ms" record and a blank line per timed stage (CudaTimer::recordSection,
   utility/CudaTimer.hpp:43-53; sections construct_bvh, refit_bvh, restr_bvh, broad_phase_cd_bvh /
   _gen / _update / _balance / _preserve, broad_phase_restr_front, broad_phase_front_ordering,
   narrow_phase) and "#cd_total_time: ms", "###: 0" (blankLine, :54-64). len = full text length;
   at most cap - 1 bytes are copied, NUL-terminated. Synchronises the stream. */
int  bvhcd_log_text(bvhcd_ctx* ctx, char* buf, int64_t cap, int64_t* len);

/* Export the current BVH to host arrays (any pointer may be NULL). n = leaves = num_tris.
   codes/prim_idx/prim_map: n; int_*: n-1 (preorder); leaf_par/leaf_mark: n; leaf_lca: n+1;
   int_box/leaf_box: 6 floats per node (minx, miny, minz, maxx, maxy, maxz). */
int  bvhcd_export_bvh(bvhcd_ctx* ctx, uint32_t* codes, int32_t* prim_idx, int32_t* prim_map,
                      int32_t* int_lc, int32_t* int_rc, int32_t* int_par, uint32_t* int_mark,
                      int32_t* int_rangex, int32_t* int_rangey,
                      int32_t* leaf_par, uint32_t* leaf_mark, int32_t* leaf_lca,
                      float* int_box, float* leaf_box);
/* Export front entries (query leaf, node) as int32 pairs; which = 0 internal, 1 external. */
int  bvhcd_export_front(bvhcd_ctx* ctx, int which, int32_t* out, int64_t cap, int64_t* n);
/* Scene box of the last BUILD (6 floats). */
int  bvhcd_export_scene_box(bvhcd_ctx* ctx, float* box6);

/* ---- learned restructure scheduler (per-subtree MLP 16->64->64->1) ---- */
#define BVHCD_MLP_IN  16
#define BVHCD_MLP_H   64
#define BVHCD_MLP_NPARAM (16*64 + 64 + 64*64 + 64 + 64 + 1)
/* Load fp32 master weights (packed W1[64][16], b1[64], W2[64][64], b2[64], W3[64], b3). */
int  bvhcd_mlp_set_weights(bvhcd_ctx* ctx, const float* params, int64_t n);
int  bvhcd_mlp_get_weights(bvhcd_ctx* ctx, float* params, int64_t n);
/* Features of the candidate subtrees of the current BVH/front: rows x 16 fp32, plus the
   preorder root of each row. */
int  bvhcd_mlp_features(bvhcd_ctx* ctx, float* feats, int32_t* roots, int64_t cap, int64_t* rows);
/* Forward pass on the device over host/device feature rows (bf16 weights, fp32 accumulate). */
int  bvhcd_mlp_forward(bvhcd_ctx* ctx, const float* feats, int64_t rows, float* out, int where);
/* bvhcd_mlp_forward over host rows without any device allocation: no counterpart in the
   reference; for callers that cannot allocate device memory (tight serving loops, harnesses
   that forbid it). Rows are staged through a static device buffer of the library (one per GPU,
   held until the process ends) instead of buffers the library grows on demand; at most
   BVHCD_MLP_NOALLOC_MAX_ROWS rows per call (more: BVHCD_E_INVALID). Calls that overlap on one
   GPU share that buffer and race. */
#define BVHCD_MLP_NOALLOC_MAX_ROWS 16384
int  bvhcd_mlp_forward_noalloc(bvhcd_ctx* ctx, const float* feats, int64_t rows, float* out);
/* One training step (fp32 master weights, bf16 forward, MSE loss, Adam) on device; returns loss. */
int  bvhcd_mlp_train_step(bvhcd_ctx* ctx, const float* feats, const float* targets, int64_t rows,
                          float lr, float* loss, int where);
/* Gradients of the MSE loss for one batch (fp32), for checking against a CPU reference. */
int  bvhcd_mlp_grad(bvhcd_ctx* ctx, const float* feats, const float* targets, int64_t rows,
                    float* grad, float* loss, int where);
/* Score threshold of the learned selection: a candidate subtree is restructured when the
   MLP's predicted gain exceeds it (default 0). */
int  bvhcd_mlp_set_threshold(bvhcd_ctx* ctx, float threshold);
/* Cost gate of the learned selection (no counterpart in the reference; the reference's
   heuristic has no cost model): a selection is applied only when its predicted saving over the
   decision interval, in front entries, exceeds repair_cost x (front entries), the KEEP work the
   incremental front repair after a rebuild costs. Decisions that change nothing (rejected or
   empty selections) space the next ones out, up to 8 bvh cycles. 0 turns the gate off;
   default 4. */
int  bvhcd_mlp_set_repair_cost(bvhcd_ctx* ctx, float repair_cost);
/* Restructure the given subtrees (preorder internal-node ids with disjoint leaf ranges) of the
   current BVH now, refit, and invalidate the front (the next step must GENERATE). The
   explicit-root form of LBvhFixedDeformable::restructure (collision/lbvh/LBvhFixedDeformable.cu:210 onward,
   whose roots come from the front's quality check); it labels training data for the learned
   scheduler. */
int  bvhcd_restructure(bvhcd_ctx* ctx, const int32_t* roots, int64_t k);

/* ---- inter-object front (BvttFrontLooseInter + LBvhRigid,
        collision/bvtt_front/BvttFrontLooseInter.{h,cu}, collision/lbvh/LBvhRigid.{h,cu}) ----
   A second object's triangles queried against a context's tree: the second object (the
   reference's rigid model: fixed topology; BUILD orders its triangles by Morton code, REFIT
   updates their boxes) gets a tree of its own whose sorted leaves are the queries, each walking
   the whole tree of the queried context, with a loose front of (query, node) entries maintained
   by the same operations as the self-collision front. Every triangle of the second object is a
   query (the reference's generate kernel skips the last one, BvttFrontGenerateKernels.cu:95).
   The queried context must have been stepped for the frame first; its tree is read, not
   changed, and a change of its topology makes the next inter step regenerate the front. Work is
   ordered on the queried context's stream. */
typedef struct bvhcd_inter bvhcd_inter;
/* cfg: max_tris / max_verts of the second object, int / ext front and pair capacities (0 =
   48 T + 2^20, 24 T + 2^20, 24 T + 2^20 of its T), narrow_phase; other fields are ignored. */
int  bvhcd_inter_create(bvhcd_ctx* queried, const bvhcd_config* cfg, bvhcd_inter** out);
int  bvhcd_inter_destroy(bvhcd_inter* x);
const char* bvhcd_inter_last_error(const bvhcd_inter* x);
int  bvhcd_inter_set_topology(bvhcd_inter* x, const int32_t* faces, int64_t num_tris, int64_t num_verts, int where);
int  bvhcd_inter_update_vertices(bvhcd_inter* x, const float* xyz, int64_t num_verts, int where);
/* rigid_op: BVHCD_BVH_BUILD or BVHCD_BVH_REFIT of the second object's tree; front_op:
   PURE_BVH_CD, GENERATE, UPDATE, REORDER or KEEP (BvttFrontLooseInter::maintain, :69-104). */
int  bvhcd_inter_step(bvhcd_inter* x, int rigid_op, int front_op);
/* Pairs (triangle of the queried context, triangle of the second object); filtered = after the
   triangle test, which takes the queried context's triangle as P. */
int  bvhcd_inter_get_pairs(bvhcd_inter* x, int32_t* out, int64_t cap_pairs, int64_t* num_pairs, int where,
                           int filtered);
/* Front entries (query = sorted leaf of the second object, node of the queried tree). */
int  bvhcd_inter_export_front(bvhcd_inter* x, int which, int32_t* out, int64_t cap, int64_t* n);
/* out[6]: internal front, external front, broad-phase pairs, narrow-phase pairs, node tests,
   last front op. */
int  bvhcd_inter_stats(bvhcd_inter* x, int64_t* out);
/* The second object's sorted codes and sorted slot -> triangle map. */
int  bvhcd_inter_export_rigid(bvhcd_inter* x, uint32_t* codes, int32_t* prim_idx);

/* ---- batched scenes (many scenes per launch; one hipGraph per frame) ----
   No counterpart object in the reference, which runs one scene per LBvh / BvttFront pair
   (Frameworks/CudaProject/CudaProjectApp.cpp:28-34); each scene here gets exactly the pairs a single-scene context
   would report for it (PURE_BVH_CD traversal + narrow phase). */
typedef struct bvhcd_batch bvhcd_batch;
/* cfg: device, stream and pair_cap are used (0 = derived from free HBM). */
int  bvhcd_batch_create(const bvhcd_config* cfg, int32_t num_scenes, const int64_t* tris_per_scene,
                        const int64_t* verts_per_scene, bvhcd_batch** out);
int  bvhcd_batch_destroy(bvhcd_batch* b);
/* faces: concatenated per-scene faces with scene-local vertex ids. */
int  bvhcd_batch_set_topology(bvhcd_batch* b, const int32_t* faces, int where);
/* xyz: concatenated per-scene vertices; host memory may be reused when the call returns. */
int  bvhcd_batch_update_vertices(bvhcd_batch* b, const float* xyz, int where);
/* One frame for every scene: bvh_op BUILD or REFIT, then the persistent self-CD traversal,
   the narrow phase and the per-scene grouping of the hits. use_graph = 1: the frame is captured
   in a hipGraph on first use (one graph per bvh_op) and replayed afterwards. */
int  bvhcd_batch_step(bvhcd_batch* b, int bvh_op, int use_graph);
/* Pairs of scene s (scene-local triangle ids, a < b), narrow-phase filtered; out = NULL only
   returns the count. Synchronises the stream. */
int  bvhcd_batch_get_pairs(bvhcd_batch* b, int32_t scene, int32_t* out, int64_t cap, int64_t* n);
int  bvhcd_batch_pair_counts(bvhcd_batch* b, int64_t* counts /* num_scenes */);
const char* bvhcd_batch_last_error(const bvhcd_batch* b);

/* ---- autotuner (replaces CudaDevice::launchConfig / evalOptimalBlockSize, CudaDevice.cu:60-94) ----
   Block sizes are fixed at 256 (wave64 x 4); the launch shapes of the grid-stride and persistent
   kernels (grid size, for KEEP the entries per thread per loop iteration, and for KEEP and the
   narrow phase the workgroup size: 128, 256 or 512 threads) are tuned by timing candidates on the
   current data, per kernel and size class ceil(log2 n), in a process-wide table used by every
   later launch. */
#define BVHCD_TUNE_KEEP      0
#define BVHCD_TUNE_NARROW    1
#define BVHCD_TUNE_BATCH_CD  2
/* KEEP and the narrow phase of a context with a built BVH and front; leaves the current front's
   pairs in place. */
int  bvhcd_autotune(bvhcd_ctx* ctx);
/* Sets the shape of one kernel for the size class of n (n = the scene's triangles for KEEP and
   the batched traversal, the context's triangle capacity for the narrow phase): grid in [1, 2^24),
   unroll 1, 2 or 4 for KEEP, 1 elsewhere. Stands where CudaDevice::launchConfig
   (system/CudaDevice/CudaDevice.cu:85-94) fixes a kernel's configuration by name. */
int  bvhcd_autotune_set(int kernel, int64_t n, uint32_t grid, uint32_t unroll);
/* The same with the workgroup size: 128, 256 or 512 threads for KEEP and the narrow phase, 256
   for the batched traversal (CudaDevice::evalOptimalBlockSize, CudaDevice.cu:60-83, picks a
   block size per kernel; bvhcd_autotune_set keeps 256). */
int  bvhcd_autotune_set_block(int kernel, int64_t n, uint32_t grid, uint32_t unroll, uint32_t block);
/* The persistent batched traversal; drops the captured frame graphs (re-captured on next use). */
int  bvhcd_batch_autotune(bvhcd_batch* b);
/* Text table of every timed candidate (kernel, size class, grid, ms, chosen); returns its
   length (copies at most cap - 1 bytes plus a terminator). */
int  bvhcd_autotune_report(char* buf, int64_t cap);

/* ---- seeded synthetic scenes (host) ---- */
#define BVHCD_SCENE_TWO_SPHERES  0   /* two jittered icospheres approaching each other */
#define BVHCD_SCENE_CLOTH        1   /* n x n folding sheet */
#define BVHCD_SCENE_LAYERS       2   /* stacked folding sheets */
#define BVHCD_SCENE_SHEET_SPHERE 3   /* folding sheet + rigid icosphere on a straight path */

typedef struct bvhcd_scene_params {
    uint64_t seed;
    int32_t  n;          /* grid vertices per side (cloth kinds) */
    int32_t  frames;     /* sequence length F */
    int32_t  layers;     /* stacked sheets (LAYERS) */
    int32_t  subdiv;     /* icosphere subdivision level */
    float    amplitude;  /* max fold amplitude */
    float    compress;   /* lateral compression of x at the last frame */
    float    sigma;      /* per-vertex jitter std (TWO_SPHERES) */
    float    radius;     /* icosphere radius */
    float    spacing;    /* layer spacing (LAYERS) or sphere centre gap (TWO_SPHERES) */
    float    speed;      /* approach speed per frame (TWO_SPHERES) */
    float    fold;       /* cloth kinds: fold angle at the last frame in radians (0 = no fold) */
    float    fold_radius;/* cloth kinds: radius of the fold's bend (<= 0: 0.05) */
} bvhcd_scene_params;

typedef struct bvhcd_scene bvhcd_scene;
int     bvhcd_scene_create(int kind, const bvhcd_scene_params* p, bvhcd_scene** out);
void    bvhcd_scene_destroy(bvhcd_scene* s);
int64_t bvhcd_scene_num_tris(const bvhcd_scene* s);
int64_t bvhcd_scene_num_verts(const bvhcd_scene* s);
int     bvhcd_scene_faces(const bvhcd_scene* s, int32_t* faces);
/* Vertices of frame f (0 <= f < frames) into xyz (V x 3 floats); threads <= 0 = all cores. */
int     bvhcd_scene_frame(const bvhcd_scene* s, int32_t frame, float* xyz, int32_t threads);

/* ---- mesh files (host) ----
   One mesh file per frame, as the reference's Scene::update / addMesh read them through assimp
   (world/Scene.cpp:14-42): every mesh of the file concatenated with vertex-id offsets. Formats:
   Wavefront OBJ (v / f lines, 1-based or negative indices, polygons split into fans) and PLY
   (ascii, binary little / big endian; vertex x y z of any scalar type, face vertex_indices
   list). Coordinates are stored as float. */
typedef struct bvhcd_mesh bvhcd_mesh;
int     bvhcd_mesh_load(const char* path, bvhcd_mesh** out);
void    bvhcd_mesh_destroy(bvhcd_mesh* m);
int64_t bvhcd_mesh_num_verts(const bvhcd_mesh* m);
int64_t bvhcd_mesh_num_tris(const bvhcd_mesh* m);
int     bvhcd_mesh_faces(const bvhcd_mesh* m, int32_t* faces);   /* num_tris x 3 */
int     bvhcd_mesh_verts(const bvhcd_mesh* m, float* xyz);       /* num_verts x 3 */
/* Path of frame `index` of a sequence named by a printf pattern with one integer conversion
   (Benchmark's inputFileFormat, setting/CDBenchmarkSettings.cpp:17-60, e.g. "funnel/%03d.ply"). */
int     bvhcd_frame_path(const char* pattern, int32_t index, char* out, int64_t cap);

#ifdef __cplusplus
}
#endif
#endif /* BVHCD_H_ */
