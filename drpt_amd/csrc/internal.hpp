// Internal declarations shared by the HIP translation units of libdrpt_hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <map>
#include <string>
#include <vector>

#include "../../include/drpt_hip.h"

namespace drpt {

void set_error(const std::string& msg);

#define DRPT_HIP(x)                                                             \
  do {                                                                          \
    hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess) {                                                     \
      drpt::set_error(std::string(#x) + " failed: " + hipGetErrorString(e_) +   \
                      " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")");  \
      return DRPT_ERR_HIP;                                                      \
    }                                                                           \
  } while (0)

#define DRPT_CHECK(x)          \
  do {                         \
    int rc_ = (x);             \
    if (rc_ != DRPT_OK) return rc_; \
  } while (0)

// pinned, device-mapped host words (host: the host's address, dev: the
// device's); grown by pinned_words, freed by drpt_destroy
struct PinnedWords {
  uint32_t* host = nullptr;
  uint32_t* dev = nullptr;
  int cap = 0;
};

}  // namespace drpt

// Device workspace: named buffers, grown on demand and kept for the context's
// lifetime so repeated builds of one shape allocate nothing.
struct drpt_ctx {
  int device = 0;
  int num_cus = 256;
  int max_lds = 65536;      // LDS bytes a workgroup may use
  bool split_lds_big = false;  // split kernels cleared for more than 64 KB of dynamic LDS
  hipStream_t stream = nullptr;
  std::map<std::string, std::pair<void*, size_t>> bufs;
  hipEvent_t ev[12] = {};
  hipEvent_t evp[4] = {};           // stage boundaries inside the leaf k-NN (drpt_last_phases)
  std::vector<hipEvent_t> ev_pass;  // per-pass timing events (grown on demand)
  drpt_phases last_phases{};        // stage times of the last drpt_build_graph
  // captured launch sequences (hipGraph), replayed while their key (sizes,
  // parameters and every buffer address they touch) is unchanged
  struct Graph {
    std::string key;
    hipGraphExec_t exec = nullptr;
  };
  std::map<std::string, Graph> graphs;
  bool use_graphs = true;  // environment DRPT_GRAPHS=0: launch eagerly
  // environment DRPT_INT8=0: the LDS-resident passes keep their bf16
  // instances for integer data in [0, 255] as well
  bool use_int8 = true;
  // pinned host copy of a build's column maxima (checked for non-finite
  // values while the splits run) and the event that marks it copied
  // pinned host words written by k_words_to_host: the projection's column
  // maxima, and the counters a low-precision build reads back at its end
  drpt::PinnedWords pin_cm, pin_st;
  hipEvent_t ev_cm = nullptr;
  // staging of the projection's term arrays, and the event after its copy
  drpt::PinnedWords pin_up;
  hipEvent_t ev_up = nullptr;
  // environment DRPT_ROW_EXACT_BATCH=r: at most r rows per batch of the exact
  // row path (its scratch is otherwise capped by bytes)
  int64_t row_exact_batch = 0;
  unsigned long long pool_leaf_lds_set = 0;  // LDS-resident pass instances cleared for their dynamic LDS
  std::map<const void*, int> fn_lds;         // dynamic LDS bytes a kernel has been cleared for
  // work-item plans found by drpt_autotune, by workload shape (ItemPlan)
  std::map<std::string, std::vector<int>> plans;
};

namespace drpt {

int ws_get(drpt_ctx* ctx, const char* name, size_t bytes, void** out);
template <typename T>
int ws(drpt_ctx* ctx, const char* name, size_t count, T** out) {
  void* p = nullptr;
  int rc = ws_get(ctx, name, count * sizeof(T) + 16, &p);
  *out = (T*)p;
  return rc;
}

// ---- host-side reference pieces (host_ref.cpp, built with g++) -------------
void host_build_projection(int d, int L, float density, uint32_t seed, float* R);
// positions of a reference local tree with all-equal keys (drpt_local.cpp:121-161)
void host_local_positions(int m, int ld, int32_t* out);
void host_local_sizes(int m, int ld, int32_t* out);
int host_local_depth(int64_t gmin, int nn, int offset);

// ---- launchers (one per .hip unit) ------------------------------------------
// util.hip: capture-safe fills (a kernel, not hipMemsetAsync)
int fill_u32(drpt_ctx* ctx, uint32_t* p, uint32_t v, size_t count);
// fills of three word ranges (p_i[0 .. c_i) = v_i) in one launch (capture-safe)
int fill3_u32(drpt_ctx* ctx, uint32_t* p0, uint32_t v0, size_t c0, uint32_t* p1, uint32_t v1,
              size_t c1, uint32_t* p2 = nullptr, uint32_t v2 = 0, size_t c2 = 0);
// pw holds at least count words
int pinned_words(drpt_ctx* ctx, PinnedWords& pw, int count);
// s0[0..c0), s1[0..c1), ... into dst (a PinnedWords dev pointer), on the
// context's stream
int words_to_host(drpt_ctx* ctx, uint32_t* dst, const uint32_t* s0, int c0,
                  const uint32_t* s1 = nullptr, int c1 = 0, const uint32_t* s2 = nullptr,
                  int c2 = 0, const uint32_t* s3 = nullptr, int c3 = 0);
// *acc += p[0] + ... + p[count - 1] (stream-ordered)
int sum_u32(drpt_ctx* ctx, const uint32_t* p, size_t count, unsigned long long* acc);
// Issues record() on the context's stream: eagerly when graphs are off, else
// captured into a hipGraph the first time `key` is seen under `name` and
// replayed from then on. record() may only launch kernels on ctx->stream (no
// allocation, no synchronisation, no host reads of device data).
int run_graph(drpt_ctx* ctx, const char* name, const std::string& key,
              const std::function<int()>& record);

// project.hip
// (rowidx / w: R's compressed columns as uint16 / float; gk / ridx32 / wd the
// same terms in groups of 4 columns (columns gcol[4g .. 4g+3], -1: none),
// interleaved and padded to multiples of 4 terms with row d and weight 0, for
// the wave-uniform kernel; may be null)
int launch_project(drpt_ctx* ctx, const void* X, int dtype, int64_t n, int d, const int32_t* colptr,
                   const uint16_t* rowidx, const float* w, const int32_t* gk, const int32_t* gcol,
                   const int32_t* ridx32, const double* wd, int L, int nnz, float* P,
                   uint32_t* colmax);

// split.hip
// device address of the allocation-free entry points' scratch array on
// `device` (current), looked up once per device by drpt_create
int noalloc_setup(int device);
int grow_trees_impl(drpt_ctx* ctx, const float* P, const uint32_t* colmax, int64_t n, int T, int D,
                    int mode, int leaf_max, uint32_t* node);

// group.hip / leafknn.hip
struct LeafItem {   // one work item: a block of rows of one leaf (or the whole leaf)
  int32_t tree;
  int32_t size;     // leaf size m
  int64_t start;    // offset of the leaf in perm[tree]
  int32_t row0;     // first row of this block within the leaf
  int32_t rows;     // rows of the block (row0 + rows <= m)
};

// Work-item layout by leaf size: leaves of bucket b (sizes <= 64, <= 256,
// <= 1024, larger) are cut into items of rpi[b] rows, and the low-precision
// Gram passes over them use the LDS tile ring (ring[b] = 1, where the row
// length allows it), the register double buffer (0) or, for the pooled filter
// pass, the whole leaf resident in LDS (2; leaves too large for it, and other
// passes, take the ring); tree_group = trees per
// pooled filter pass (0: the rule of drpt_params.tree_group = 0); first_trees =
// trees whose widened candidates give the first bounds (0: 4); rerank = their
// best approximate candidates whose exact distances give those bounds, in
// [K, 2K] (0: 2K; -1: the bounds come from upper bounds of the approximate
// distances, no exact distance). Chosen by drpt_autotune; the default is whole
// leaves, the ring and the rules.
constexpr int kSizeBuckets = 4;
constexpr int kPlanInts = 2 * kSizeBuckets + 3;  // rpi, ring, tree_group, first_trees, rerank
struct ItemPlan {
  int rpi[kSizeBuckets];
  int ring[kSizeBuckets];
  int tree_group;
  int first_trees;
  int rerank;
};
__host__ __device__ inline int size_bucket(int m) {
  return m <= 64 ? 0 : m <= 256 ? 1 : m <= 1024 ? 2 : 3;
}
// [lo, hi] leaf sizes of bucket b
inline void bucket_sizes(int b, int* lo, int* hi) {
  static const int lim[kSizeBuckets + 1] = {0, 64, 256, 1024, 1 << 30};
  *lo = lim[b] + 1;
  *hi = lim[b + 1];
}
inline ItemPlan uniform_plan(int rpi) {
  ItemPlan p;
  for (int b = 0; b < kSizeBuckets; ++b) {
    p.rpi[b] = rpi;
    p.ring[b] = 1;
  }
  p.tree_group = 0;
  p.first_trees = 0;
  p.rerank = 0;
  return p;
}

int launch_count_leaves(drpt_ctx* ctx, const uint32_t* heap, int64_t n, int T, int D, uint32_t* hcount);
int launch_scatter_leaves(drpt_ctx* ctx, const uint32_t* heap, int64_t n, int T, int D,
                          const int64_t* hoff, uint32_t* cursor, int32_t* perm);
// device-side grouping for leaf-max builds (hcount: n points per tree): heap
// offsets, first work item of each tree (base[T + 1]), item ranges (trees <
// first_trees / the others / all) and the work items (blocks of
// plan.rpi[bucket] rows); stats = {sum m^2, max m, leaves}
int launch_device_items(drpt_ctx* ctx, const uint32_t* hcount, int64_t n, int T, int D,
                        const ItemPlan& plan, int first_trees, int64_t* hoff, int64_t* base,
                        int64_t* range, LeafItem* items, unsigned long long* stats);
// Xp[i] = X[perm[i]] (rows of rowbytes bytes) and pos[perm[i]] = i, i < n
int launch_permute_rows(drpt_ctx* ctx, const void* X, size_t rowbytes, const int32_t* perm,
                        int64_t n, void* Xp, int32_t* pos);
// range[0] = base[t0], range[1] = base[t1]: the work items of trees [t0, t1)
int launch_tree_range(drpt_ctx* ctx, const int64_t* base, int t0, int t1, int64_t* range);
// out[0, out_range[1]) = the items of [range[0], range[1]) with leaf sizes in
// [lo, hi] (out_range[0] = 0); max_items bounds the input range and out's size
int launch_select_items(drpt_ctx* ctx, const LeafItem* items, const int64_t* range, int lo, int hi,
                        int64_t max_items, LeafItem* out, int64_t* out_range);
int launch_sort_segments(drpt_ctx* ctx, int32_t* perm, int64_t n, const int64_t* seg_start,
                         const int32_t* seg_size, const int32_t* seg_tree, int nseg, int max_seg);
// sorts each segment of perm by (LP[seg_tree * ld + level][point], point):
// one level of the real-key local tree
int launch_sort_keyed(drpt_ctx* ctx, int32_t* perm, int64_t n, const float* LP, int ld, int level,
                      const int64_t* seg_start, const int32_t* seg_size, const int32_t* seg_tree,
                      int nseg, int max_seg);
int launch_apply_local(drpt_ctx* ctx, const int32_t* perm_in, int32_t* perm_out, int64_t n,
                       const int64_t* seg_start, const int32_t* seg_tree, const int64_t* seg_tbl,
                       const int32_t* seg_size, int nseg, const int32_t* tables);

// quantised rows (bf16: n x d halves; fp8: n x d bytes with a power-of-two
// scale 2^pexp[i] per row), nbe[i] = (|x~|^2, |x~ - x|) of the dequantised rows;
// bf16 of float X with notu8 given (zeroed by the caller): also the int8 copy
// xi8 = x - 128 (n x d bytes) and *notu8 = 1 when some value is not an integer
// in [0, 255]
int launch_prep_lowp(drpt_ctx* ctx, const void* X, int dtype, int64_t n, int d, int precision,
                     void* Xq, float2* nbe, int8_t* pexp, uint8_t* xi8 = nullptr,
                     int32_t* notu8 = nullptr);
// *lossy = 1 when some row's quantisation error (nbe[i].y) is not 0 (the caller zeroes it)
int launch_lossy_flag(drpt_ctx* ctx, const float2* nbe, int64_t n, int32_t* lossy);
// leaf-ordered |x~|^2, |x~ - x| and (fp8) scale exponents: (lnbT[i], lerrT[i])
// = nbe[perm[i]], lexpT[i] = pexp[perm[i]] (pexp may be null)
int launch_leaf_gather(drpt_ctx* ctx, const int32_t* perm, int64_t total, const float2* nbe,
                       const int8_t* pexp, float* lnbT, float* lerrT, int8_t* lexpT,
                       const int32_t* gate = nullptr);
// FP32 exact per-tree lists for items[range[0] .. range[1]) (range in device memory)
int launch_leaf_knn(drpt_ctx* ctx, const drpt_params* p, const void* X, const int32_t* perm,
                    const LeafItem* items, const int64_t* range, int32_t* cand_idx,
                    float* cand_dist);
// ---- low-precision (BF16 / FP8) leaf k-NN, leafknn.hip -----------------------
// modes of launch_lowp_knn
enum { kModeTree = 0, kModeCand = 1, kModePool = 2 };

struct KnnArgs {
  const void* X;
  const void* Xq;        // quantised rows (bf16 or fp8 e4m3), n x d
  int precision;         // DRPT_BF16 | DRPT_FP8
  const int8_t* lexpT;   // fp8: leaf-ordered scale exponents, [T][n]
  const float* lnbT;     // leaf-ordered |x~|^2, [T][n]
  const float* lerrT;    // leaf-ordered |x~ - x|, [T][n]
  const int32_t* perm;   // [T][n] points in leaf order
  const LeafItem* items;
  const int64_t* range;  // items [range[0], range[1])
  int64_t n;
  int d, K;
  int size_lo, size_hi;  // items of leaves with sizes outside [lo, hi] are skipped
  int ring;              // 1: LDS tile ring where the row length allows it; 0: registers
  unsigned* queue;       // ticket counters (set by launch_lowp_knn)
  // kModeTree: per-tree lists [T][n][K]; kModePool: pool ids [n][pool_cap]
  int32_t* out_idx;
  // kModeCand: keys approx_bits << 32 | id, [tree - tree0][n][k']
  uint64_t* out_key;
  float* out_dist;
  int tree0;
  // kModeTree: rows redone by k_knn_fallback
  int32_t* fallback;
  int32_t* nfallback;
  int64_t fb_cap;
  // kModePool
  const float* boundT;   // leaf-ordered K-th distance bound, [T][n]
  const int32_t* known;  // [n][K] ids already evaluated (-1 = none)
  int32_t* pool_cnt;     // [n] entries of each row's pool (zeroed by the caller)
  int pool_cap;          // slots per row
  int32_t* ovf_flag;     // [n]
  int32_t* ovf_list;     // rows whose pool overflowed
  int32_t* novf;
  // optional int8 copy of X (x - 128, n x d) and a device flag, 0 when every
  // value of X is an integer in [0, 255]: the LDS-resident passes then run
  // their int8 instances (and their bf16 ones return at once)
  const void* Xi8;
  const int32_t* notu8;
  int64_t item_cap;      // capacity of the item list (items gathered by size need as much)
  // optional: lnbT / lerrT were not written when *lnb_gate == 0; the streaming
  // kernel then reads (|x~|^2, |x~ - x|) from nbe by id
  const float2* nbe;
  const int32_t* lnb_gate;
};
// kModeTree: exact per-tree lists (bound pass + filter pass); kModeCand: the
// k' widened candidates of each row (k' = kprime_eff(p)); kModePool: filter
// against an external bound, survivors appended to per-row pools
int launch_lowp_knn(drpt_ctx* ctx, const drpt_params* p, int mode, KnnArgs a);
int lowp_slots(const drpt_params* p);  // k' actually used (a multiple of 4)
// leafpool.hip: the pool and candidate passes with each leaf resident in LDS
// (work-item plan flag 2): largest leaf it stages at dimension d in a mode
// (0: not applicable), and the launch (mode kModePool or kModeCand with kp = 4 TK
// widened candidates, TK in {3, 4, 6}; BF16 rows) over the items of a.range
// with sizes in [a.size_lo, a.size_hi]
int pool_leaf_rows(drpt_ctx* ctx, int d, int mode);
// the int8 instances apply at dimension d (given an int8 copy and its flag)
bool pool_leaf_int8(int d);
int launch_pool_leaf(drpt_ctx* ctx, int mode, int kp, const KnnArgs& a);

struct RowMergeArgs {
  const void* X;
  int64_t n;
  int d, K;
  // candidates (duplicates, -1 and the row itself allowed) from ntsrc sources
  // of S slots each: slot s of source t of a row is at (t * src_tree_stride +
  // row * S + s), valid for s < cnt_tree[t * n + row] (all S when cnt_tree is
  // null); as ids (src) or as keys approx_bits << 32 | id (src_key)
  const int32_t* src;
  const uint64_t* src_key;
  int64_t src_tree_stride;
  int S, ntsrc;
  const int32_t* cnt_tree;
  int rerank;                // src_key: exact D for the rerank best approximate candidates only
  const int32_t* known_idx;  // optional [n][K] list (sorted, -1 padded) merged in
  const float* known_dist;
  const int32_t* skip_flag;  // optional [n]: rows left to the exact row path
  const int32_t* order;      // optional permutation of the rows (processing order)
  const int32_t* xpos;       // optional: row i of the data is X[xpos[i]] (X a permuted copy)
  // optional bf16 copy of X (row i at Xb + i * d) and a device flag, 0 when
  // every row of it equals its X row exactly: the exact distances then read
  // the bf16 rows (same values, half the bytes)
  const uint16_t* Xb;
  const int32_t* lossy;
  // optional int8 copy of X (x - 128) and a device flag, 0 when every value
  // of X is an integer in [0, 255] (the copy then holds them all): read before Xb
  const uint8_t* Xi8;
  const int32_t* notu8;
  int32_t* out_idx;          // [n][K]
  float* out_dist;
  float* out_bound;          // optional [n]: out_dist[row][K - 1] (the row's bound)
  // optional, one source with more than 256 slots: rows with at most 256
  // candidates are merged by a light pass (fewer registers per wave, more waves
  // in flight), the others listed in heavy[0 .. *nheavy) for the full pass
  int32_t* heavy;            // [n]
  int32_t* nheavy;           // [1]
  const int32_t* nrows_dev;  // kernel: rows to process read on the device (order = heavy)
};
int launch_row_merge(drpt_ctx* ctx, int dtype, RowMergeArgs a);
// out_bound[row] = K-th smallest upper bound of the exact distance over the
// distinct ids among a row's candidate keys (slot s of source t of a row at
// t * tree_stride + row * S + s), from the keys' approximate distances and
// nbe = (|x~|^2, |x~ - x|) of the quantisation they were computed in
int launch_row_bound(drpt_ctx* ctx, const uint64_t* src_key, int64_t tree_stride, int S, int ntsrc,
                     int64_t n, int d, int K, const float2* nbe, float* out_bound);
// rows listed in ovf_list: exact K best over every leafmate in every tree
int launch_row_exact(drpt_ctx* ctx, const drpt_params* p, const void* X, const int32_t* perm,
                     const LeafItem* items, const int64_t* range, const int32_t* ovf_list,
                     const int32_t* novf, int64_t nrows_max, int32_t* out_idx, float* out_dist);
// leaf-ordered gather of a per-point value: outT[i] = src[perm[i] * stride + offset]
int launch_leaf_gather_val(drpt_ctx* ctx, const int32_t* perm, int64_t total, const float* src,
                           int stride, int offset, float* outT);

int launch_fallback(drpt_ctx* ctx, const drpt_params* p, const void* X, const int32_t* perm,
                    const LeafItem* items, const int32_t* fallback, const int32_t* nfallback,
                    int64_t fallback_cap, int32_t* cand_idx, float* cand_dist);

// merge.hip
int launch_merge(drpt_ctx* ctx, int64_t n, int T, int K, const int32_t* cand_idx,
                 const float* cand_dist, int32_t* out_idx, float* out_dist);

}  // namespace drpt
