/*
 * drpt_hip.h — C ABI of the MI355X-native DRPT k-NN graph builder.
 *
 * The reference (HipGraph/DRPT) has no FFI: its hot path sits behind the C++
 * class API of libmdrpt.so, called from cpp/src/main.cpp:198-217:
 *   MDRPT(int ntrees, int tree_depth, double tree_depth_ratio, ...)   mdrpt.hpp:102-105
 *   void grow_trees(vector<vector<float>>&, float density, bool, int nn, ofstream&)
 *                                                                      mdrpt.hpp:107-108
 *   map<int, vector<DataPoint>> gather_nns(int nn, ofstream&)          mdrpt.hpp:110
 * with kernel seams in MathOp (cpp/src/drpt/math/math_operations.hpp:18-46).
 * Each entry point below names the reference interface it replaces.
 *
 * Conventions: plain C types; all data pointers marked (device) are HIP
 * device pointers owned by the caller (outputs are caller-allocated); calls
 * are ordered on the context's stream; status codes are 0 or negative and
 * drpt_last_error() returns a thread-local message. No exceptions cross the
 * ABI. One context per host thread.
 */
#ifndef DRPT_HIP_H
#define DRPT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DRPT_OK 0
#define DRPT_ERR_ARG (-1)         /* invalid argument (reference: exit(1) in main.cpp:99-127) */
#define DRPT_ERR_HIP (-2)         /* HIP runtime error */
#define DRPT_ERR_NOMEM (-3)       /* device allocation failed */
#define DRPT_ERR_DEPTH (-4)       /* depth out of range (reference throws std::out_of_range,
                                     drpt_global.cpp:155-158, drpt_local.cpp:84-87) */
#define DRPT_ERR_UNSUPPORTED (-5) /* configuration this build does not implement */
#define DRPT_ERR_NONFINITE (-6)   /* non-finite projected values */

typedef struct drpt_ctx drpt_ctx;

enum { DRPT_U8 = 0, DRPT_F32 = 1 };                     /* input element type */
enum { DRPT_FP32 = 0, DRPT_BF16 = 1, DRPT_FP8 = 2 };    /* candidate-distance precision */
enum { DRPT_LOCAL_REFERENCE = 0, DRPT_LOCAL_LEAFMAX = 1 };
enum { DRPT_TREE_FIXED = 0, DRPT_TREE_LEAFMAX = 1 };

typedef struct {
  int64_t n;            /* points (-data-set-size) */
  int32_t d;            /* dimension (-dimension) */
  int32_t dtype;        /* DRPT_U8 | DRPT_F32 */
  int32_t ntrees;       /* -ntrees */
  int32_t depth;        /* projection columns per tree: Dg (reference mode) or max depth */
  int32_t nn;           /* -nn: K = nn non-self neighbours returned per point (1-64;
                           BF16 / FP8 builds: 1-32) */
  int32_t local_mode;   /* DRPT_LOCAL_REFERENCE | DRPT_LOCAL_LEAFMAX */
  int32_t leaf_max;     /* leaf-max mode: nodes larger than this are split */
  int32_t local_offset; /* -local-tree-offset (reference mode, main.cpp:37) */
  int32_t precision;    /* DRPT_FP32 | DRPT_BF16 | DRPT_FP8 */
  int32_t kprime;       /* widened candidate count for BF16/FP8 (>= 2*nn; 0 = default) */
  uint32_t seed;        /* projection seed (reference draws it from random_device) */
  float density;        /* projection density (main.cpp:129-132 default 1/sqrt(d)) */
  int32_t tree_group;   /* BF16/FP8 builds: trees per pooled filter pass, each followed by a
                           merge that tightens the bound (0 = automatic) */
  int32_t locality;     /* -locality (main.cpp:39, 0/1): BF16/FP8 builds' row merges read
                           their rows from a copy of X in tree 0's leaf order (the one-GPU
                           analogue of calculate_tree_leaf_correlation, drpt_global.cpp:572-732:
                           leafmates' rows stored together); results are unchanged */
  int32_t reserved[3];
} drpt_params;

typedef struct {
  double project_ms, split_ms, group_ms, knn_ms, merge_ms, total_ms;
  double gram_flops;       /* algorithmic Gram flops, 2*d*sum(m^2) over final leaves */
  double knn_kernel_ms;    /* leaf k-NN kernel alone */
  int64_t n_leaves;        /* final leaves over all trees */
  int64_t n_fallback_rows; /* low-precision rows re-done exactly (candidate-list overflow, leaf > 2048) */
  int64_t max_leaf;        /* largest final leaf */
  int64_t reserved[4];
} drpt_timings;

/* Stage times of the last drpt_build_graph on a context (HIP events on its
 * stream). BF16 / FP8 builds: grouping, quantised rows and leaf-ordered norms,
 * candidate pass, first row merge (known lists), pooled filter passes, the
 * row merges after them, leaf-ordered bound gathers, exact row path. FP32
 * builds: grouping, the exact leaf k-NN kernel (pool_ms), the cross-tree merge
 * (pool_merge_ms) and the fallback rows (exact_ms). No counterpart in the
 * reference, which times index building and querying as two spans
 * (main.cpp:196-220). */
typedef struct {
  double group_ms, prep_ms, cand_ms, known_ms, pool_ms, pool_merge_ms, bound_ms, exact_ms;
  int64_t tree_groups;  /* pooled filter passes (trees per pass: ntrees / tree_groups, rounded up) */
  int64_t first_trees;  /* trees whose widened candidates gave the first bounds */
  int64_t pool_cap;     /* pool slots per row */
  int64_t rerank;       /* exact re-rank depth of the first bounds (-1: bounds from the
                           approximate distances alone) */
  int64_t pool_entries; /* leafmates the pooled filter passes appended, over all rows and passes */
  int64_t int8_passes;  /* 1: the LDS-resident candidate and pool passes ran on int8 rows x - 128
                           (every value of X an integer in [0, 255]; exact int32 Gram) */
  int64_t reserved[2];
} drpt_phases;

/* Thread-local message of the last failing call. */
const char* drpt_last_error(void);
int drpt_version(void);

/* Context: one HIP device + one stream + a device workspace. */
int drpt_create(drpt_ctx** ctx, int device);
int drpt_destroy(drpt_ctx* ctx);
int drpt_sync(drpt_ctx* ctx);
/* The context's hipStream_t, for callers that order their own work on it. */
void* drpt_get_stream(drpt_ctx* ctx);

/* Parameter derivation. Replaces main.cpp:27-41 (defaults), :129-132
 * (density), :141-145 (tree_depth = int(log2 n - 2)) and mdrpt.cpp:52
 * (Dg = int(tree_depth * ratio)). Fills out with the reference configuration
 * (local_mode = DRPT_LOCAL_REFERENCE, depth = Dg, precision FP32). */
int drpt_derive_params(int64_t n, int d, int ntrees, int nn, double tree_depth_ratio,
                       drpt_params* out);

/* Host: the sparse Gaussian projection matrix R[d][L] (row-major), as
 * MathOp::build_sparse_local_random_matrix (math_operations.cpp:48-76). */
int drpt_build_projection(int d, int L, float density, uint32_t seed, float* R_host);

/* Stage (a): P[c][i] = X.R (column-major, L x n, device), and per-column
 * max |P| bit patterns (device, L). Replaces MathOp::multiply_mat
 * (math_operations.cpp:15-46) as called at mdrpt.cpp:63-66. */
int drpt_project(drpt_ctx* ctx, const void* X, int dtype, int64_t n, int d, const float* R_host,
                 int L, float* P, uint32_t* colmax);

/* Stage (b): grow ntrees trees over P (device); leaf_heap[t*n + i] (device)
 * receives the heap index (root 1) of point i's leaf in tree t.
 * tree_mode DRPT_TREE_FIXED splits every node at levels 0..depth-2 as
 * DRPTGlobal::grow_global_tree / grow_global_subtree (drpt_global.cpp:153-352);
 * DRPT_TREE_LEAFMAX splits only nodes larger than leaf_max. */
int drpt_grow_trees(drpt_ctx* ctx, const float* P, const uint32_t* colmax, int64_t n, int ntrees,
                    int depth, int tree_mode, int leaf_max, uint32_t* leaf_heap);

/* Allocation-free variant of drpt_grow_trees: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * stream capture, serving loops, harnesses that forbid hipMalloc). Same
 * arguments and results as drpt_grow_trees (bit for bit); the per-node split
 * state lives in a 32 MiB device array of the library instead of the
 * context's workspace, so ntrees * 2^(depth-2) must be <= 196608 (else
 * DRPT_ERR_ARG), and the launches are issued eagerly on the context's stream
 * (no hipGraph of its own). The array is one per GPU and holds the last
 * call's state: calls that overlap on one GPU (several contexts or streams
 * at once) race. */
int drpt_grow_trees_noalloc(drpt_ctx* ctx, const float* P, const uint32_t* colmax, int64_t n,
                            int ntrees, int depth, int tree_mode, int leaf_max,
                            uint32_t* leaf_heap);

/* Stages local + (c): from leaf_heap (device), form the final leaves
 * (reference mode: MDRPT::grow_local_trees, mdrpt.cpp:321-400 with
 * drpt_local.cpp; leaf-max mode: the leaves as grown) and compute, per tree
 * and point, the K = p->nn nearest non-self leafmates ordered by (exact FP32
 * squared distance, index): cand_idx/cand_dist[t][i][k] (device).
 * Replaces MDRPT::calculate_nns (mdrpt.cpp:123-212) for every tree. */
int drpt_tree_knn(drpt_ctx* ctx, const drpt_params* p, const void* X, const uint32_t* leaf_heap,
                  int32_t* cand_idx, float* cand_dist, drpt_timings* timings);

/* Stage (d): merge T per-tree lists into the final K per point, ordered by
 * (distance, index) with duplicates removed. Replaces the std::merge/unique
 * of mdrpt.cpp:180-198 and the one-rank communicate_nns (mdrpt.cpp:625-630). */
int drpt_merge(drpt_ctx* ctx, int64_t n, int T, int K, const int32_t* cand_idx,
               const float* cand_dist, int32_t* out_idx, float* out_dist);

/* One-shot build: MDRPT::grow_trees + MDRPT::gather_nns (main.cpp:198-217).
 * X (device, n x d row-major, u8 or f32), out_idx/out_dist (device, n x nn)
 * hold, per point, the nn nearest non-self neighbours found and their exact
 * FP32 squared distances (-1 / +inf padding). */
int drpt_build_graph(drpt_ctx* ctx, const drpt_params* p, const void* X, int32_t* out_idx,
                     float* out_dist, drpt_timings* timings);

/* Copies the stage times of the context's last drpt_build_graph to out. */
int drpt_last_phases(drpt_ctx* ctx, drpt_phases* out);

/* Work-item autotuner (no counterpart in the reference, whose leaf loop is
 * one OpenMP loop over leaves, mdrpt.cpp:123-212). For a BF16/FP8 workload
 * (X and p as for drpt_build_graph), times the pooled filter pass of a build
 * per leaf-size bucket (sizes <= 64, <= 256, <= 1024, larger) over the rows a
 * work item takes (whole leaf, 256, 128, 64) and the Gram pipeline (LDS tile
 * ring, register double buffer, or, for whole leaves whose bf16 rows fit the
 * LDS, the leaf resident in LDS), then the whole low-precision stage per
 * trees-per-group (ntrees, 8, 6, 5, 4, 3, 2; when p->tree_group is 0), per
 * number of trees whose widened candidates give the first bounds (4, 2, 1) and
 * per number of their best approximate candidates re-ranked exactly for those
 * bounds (2 nn, 1.5 nn, nn, or none: -1, bounds from upper bounds of the
 * approximate distances). Keeps the fastest in the context and writes it to
 * plan_out[11]: rows per item of buckets 0-3, their ring flags, trees per
 * group, first trees, re-rank depth. Later drpt_build_graph calls
 * with the same n, d, dtype, precision, nn, ntrees, depth, leaf_max,
 * local_mode, kprime and tree_group use it; results are unchanged by the
 * plan. */
int drpt_autotune(drpt_ctx* ctx, const drpt_params* p, const void* X, int32_t* plan_out);

/* Sets the work-item plan of a workload shape, in drpt_autotune's plan_out
 * layout (rows per item >= 64, pipeline flags 0 (registers) / 1 (LDS tile
 * ring) / 2 (leaf resident in LDS, pooled filter pass), trees per group and first
 * trees >= 0 and re-rank depth >= -1, with 0 for the automatic rules; a depth
 * above 0 is clamped to [nn, 2 nn], -1 takes the first bounds from the
 * approximate distances): restores a plan saved from an earlier drpt_autotune. No
 * counterpart in the reference. */
int drpt_set_item_plan(drpt_ctx* ctx, const drpt_params* p, const int32_t* plan);

#ifdef __cplusplus
}
#endif

#endif /* DRPT_HIP_H */
