// C ABI of libdrpt_hip: context, workspace, parameter derivation and the
// orchestration of the four stages (projection, splits, leaf k-NN, merge).
//
// The reference drives the same sequence from MDRPT::grow_trees
// (cpp/src/drpt/algo/mdrpt.cpp:46-121) and MDRPT::gather_nns (:273-291) on
// every MPI rank; here one context drives one GPU and one stream.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>

#include "internal.hpp"

namespace {
thread_local std::string g_error;
}

namespace drpt {

void set_error(const std::string& msg) { g_error = msg; }

int ws_get(drpt_ctx* ctx, const char* name, size_t bytes, void** out) {
  auto it = ctx->bufs.find(name);
  if (it != ctx->bufs.end() && it->second.second >= bytes) {
    *out = it->second.first;
    return DRPT_OK;
  }
  if (it != ctx->bufs.end()) {
    DRPT_HIP(hipStreamSynchronize(ctx->stream));
    DRPT_HIP(hipFree(it->second.first));
    ctx->bufs.erase(it);
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    set_error(std::string("device allocation of ") + std::to_string(bytes) + " bytes for '" + name +
              "' failed: " + hipGetErrorString(e));
    return DRPT_ERR_NOMEM;
  }
  ctx->bufs[name] = {p, bytes};
  *out = p;
  return DRPT_OK;
}

namespace {

template <typename T>
int upload(drpt_ctx* ctx, const char* name, const std::vector<T>& host, T** dev) {
  DRPT_CHECK(ws(ctx, name, std::max<size_t>(host.size(), 1), dev));
  if (!host.empty())
    DRPT_HIP(hipMemcpyAsync(*dev, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice,
                            ctx->stream));
  return DRPT_OK;
}

int validate(const drpt_params* p) {
  if (!p) {
    set_error("params is null");
    return DRPT_ERR_ARG;
  }
  if (p->n < 2 || p->n > INT32_MAX) {
    set_error("n must be in [2, 2^31-1]");
    return DRPT_ERR_ARG;
  }
  if (p->d < 1 || p->d > 65535) {
    set_error("dimension must be in [1, 65535]");
    return DRPT_ERR_ARG;
  }
  if (p->dtype != DRPT_U8 && p->dtype != DRPT_F32) {
    set_error("dtype must be DRPT_U8 or DRPT_F32");
    return DRPT_ERR_ARG;
  }
  if (p->ntrees < 1) {
    set_error("no of trees should be greater than zero");  // drpt_global.cpp:162
    return DRPT_ERR_DEPTH;
  }
  if (p->depth <= 0 || p->depth > std::log2((double)p->n) + (p->local_mode == DRPT_LOCAL_LEAFMAX ? 64 : 0)) {
    set_error("depth should be in range [1,....,log2(rows)]");  // drpt_global.cpp:155-158
    return DRPT_ERR_DEPTH;
  }
  if (p->depth > 20) {
    set_error("depth > 20 is not supported");
    return DRPT_ERR_UNSUPPORTED;
  }
  if (p->nn < 1 || p->nn > 64) {
    set_error("nn must be in [1, 64]");
    return DRPT_ERR_ARG;
  }
  if (p->local_mode == DRPT_LOCAL_LEAFMAX && p->leaf_max < 1) {
    set_error("leaf_max must be >= 1 in leaf-max mode");
    return DRPT_ERR_ARG;
  }
  if (p->precision != DRPT_FP32 && p->precision != DRPT_BF16 && p->precision != DRPT_FP8) {
    set_error("precision not supported by this build");
    return DRPT_ERR_UNSUPPORTED;
  }
  if (p->precision != DRPT_FP32 && p->nn > 32) {
    // the exact row paths of the low-precision builds keep at most 32 per row
    set_error("nn > 32 needs precision FP32 (BF16 / FP8 builds keep at most 32 neighbours)");
    return DRPT_ERR_UNSUPPORTED;
  }
  if (p->tree_group < 0) {
    set_error("tree_group must be >= 0");
    return DRPT_ERR_ARG;
  }
  return DRPT_OK;
}

int kprime_of(const drpt_params* p) { return p->kprime > 0 ? p->kprime : std::min(64, 2 * p->nn + 4); }

float ms_between(hipEvent_t a, hipEvent_t b) {
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return -1.0f;
  return ms;
}

struct Leaf {
  int32_t tree;
  int32_t size;
  int64_t start;
};

}  // namespace

}  // namespace drpt

using namespace drpt;

extern "C" {

const char* drpt_last_error(void) { return g_error.c_str(); }

int drpt_version(void) { return 100; }

int drpt_create(drpt_ctx** out, int device) {
  if (!out) {
    set_error("ctx out pointer is null");
    return DRPT_ERR_ARG;
  }
  *out = nullptr;
  int count = 0;
  DRPT_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) {
    set_error("device index out of range (" + std::to_string(count) + " devices)");
    return DRPT_ERR_ARG;
  }
  DRPT_HIP(hipSetDevice(device));
  DRPT_CHECK(noalloc_setup(device));
  drpt_ctx* ctx = new drpt_ctx();
  ctx->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cus = prop.multiProcessorCount;
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess &&
      lds > 0)
    ctx->max_lds = lds;
  if (const char* g = getenv("DRPT_GRAPHS")) ctx->use_graphs = strcmp(g, "0") != 0;
  if (const char* g = getenv("DRPT_INT8")) ctx->use_int8 = strcmp(g, "0") != 0;
  if (const char* b = getenv("DRPT_ROW_EXACT_BATCH")) ctx->row_exact_batch = std::atoll(b);
  if (hipStreamCreate(&ctx->stream) != hipSuccess) {
    delete ctx;
    set_error("hipStreamCreate failed");
    return DRPT_ERR_HIP;
  }
  for (auto& e : ctx->evp) {
    if (hipEventCreate(&e) != hipSuccess) {
      (void)hipGetLastError();
      e = nullptr;
      drpt_destroy(ctx);
      set_error("hipEventCreate failed");
      return DRPT_ERR_HIP;
    }
  }
  for (auto& e : ctx->ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      (void)hipGetLastError();
      e = nullptr;
      drpt_destroy(ctx);  // the stream and the events created so far
      set_error("hipEventCreate failed");
      return DRPT_ERR_HIP;
    }
  }
  *out = ctx;
  return DRPT_OK;
}

int drpt_destroy(drpt_ctx* ctx) {
  if (!ctx) return DRPT_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& kv : ctx->graphs)
    if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
  for (auto& kv : ctx->bufs) (void)hipFree(kv.second.first);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : ctx->evp)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : ctx->ev_pass) (void)hipEventDestroy(e);
  if (ctx->ev_cm) (void)hipEventDestroy(ctx->ev_cm);
  if (ctx->pin_cm.host) (void)hipHostFree(ctx->pin_cm.host);
  if (ctx->pin_st.host) (void)hipHostFree(ctx->pin_st.host);
  if (ctx->pin_up.host) (void)hipHostFree(ctx->pin_up.host);
  if (ctx->ev_up) (void)hipEventDestroy(ctx->ev_up);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return DRPT_OK;
}

int drpt_sync(drpt_ctx* ctx) {
  if (!ctx) {
    set_error("ctx is null");
    return DRPT_ERR_ARG;
  }
  DRPT_HIP(hipStreamSynchronize(ctx->stream));
  return DRPT_OK;
}

void* drpt_get_stream(drpt_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int drpt_derive_params(int64_t n, int d, int ntrees, int nn, double tree_depth_ratio,
                       drpt_params* out) {
  if (!out || n < 2 || d < 1 || ntrees < 1 || nn < 1) {
    set_error("derive_params: invalid arguments");
    return DRPT_ERR_ARG;
  }
  std::memset(out, 0, sizeof(*out));
  out->n = n;
  out->d = d;
  out->dtype = DRPT_F32;
  out->ntrees = ntrees;
  const int tree_depth = (int)(std::log2((double)n) - 2);           // main.cpp:141-145
  out->depth = (int)(tree_depth * tree_depth_ratio);                 // mdrpt.cpp:52
  out->nn = nn;
  out->local_mode = DRPT_LOCAL_REFERENCE;
  out->leaf_max = 0;
  out->local_offset = 2;                                             // main.cpp:37
  out->precision = DRPT_FP32;
  out->kprime = 0;
  out->seed = 1;
  out->density = (float)(1.0 / std::sqrt((double)d));                // main.cpp:129-132
  return DRPT_OK;
}

int drpt_build_projection(int d, int L, float density, uint32_t seed, float* R_host) {
  if (d < 1 || L < 1 || !R_host) {
    set_error("build_projection: invalid arguments");
    return DRPT_ERR_ARG;
  }
  host_build_projection(d, L, density, seed, R_host);
  return DRPT_OK;
}

int drpt_project(drpt_ctx* ctx, const void* X, int dtype, int64_t n, int d, const float* R_host,
                 int L, float* P, uint32_t* colmax) {
  if (!ctx || !X || !R_host || !P || !colmax || n < 1 || d < 1 || d > 65535 || L < 1) {
    set_error("project: invalid arguments");
    return DRPT_ERR_ARG;
  }
  DRPT_HIP(hipSetDevice(ctx->device));
  std::vector<int32_t> colptr(L + 1, 0), ridx32;
  std::vector<uint16_t> rowidx;
  std::vector<float> w;
  std::vector<double> wd;
  for (int c = 0; c < L; ++c) {
    colptr[c] = (int32_t)rowidx.size();
    for (int j = 0; j < d; ++j) {
      const float v = R_host[(size_t)j * L + c];
      if (v != 0.0f) {
        rowidx.push_back((uint16_t)j);
        w.push_back(v);
      }
    }
  }
  colptr[L] = (int32_t)rowidx.size();
  // the same terms grouped by 4 columns of similar non-zero counts (columns
  // in order of count), interleaved, padded per group to a multiple of 4 terms
  // with (row d: a zero slot, weight 0); gcol[4g + u] = column of slot u
  const int ngroups = (L + 3) / 4;
  std::vector<int32_t> order(L), gk(ngroups + 1, 0), gcol(4 * (size_t)ngroups, -1);
  for (int c = 0; c < L; ++c) order[c] = c;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return colptr[a + 1] - colptr[a] > colptr[b + 1] - colptr[b];
  });
  ridx32.clear();
  wd.clear();
  for (int g = 0; g < ngroups; ++g) {
    int kmax = 0;
    for (int u = 0; u < 4 && 4 * g + u < L; ++u) {
      const int c = order[4 * g + u];
      gcol[4 * (size_t)g + u] = c;
      kmax = std::max(kmax, colptr[c + 1] - colptr[c]);
    }
    kmax = (kmax + 3) & ~3;
    gk[g + 1] = gk[g] + kmax;
    for (int k = 0; k < kmax; ++k)
      for (int u = 0; u < 4; ++u) {
        const int c = gcol[4 * (size_t)g + u];
        const bool has = c >= 0 && k < colptr[c + 1] - colptr[c];
        ridx32.push_back(has ? (int32_t)rowidx[colptr[c] + k] : d);
        wd.push_back(has ? (double)w[colptr[c] + k] : 0.0);
      }
  }
  // the seven term arrays at 256-byte offsets in one pinned staging buffer,
  // sent with one copy; the staging buffer is refilled only once the
  // previous call's copy out of it has completed
  const std::pair<const void*, size_t> parts[7] = {
      {colptr.data(), colptr.size() * 4}, {rowidx.data(), rowidx.size() * 2},
      {w.data(), w.size() * 4},           {gk.data(), gk.size() * 4},
      {gcol.data(), gcol.size() * 4},     {ridx32.data(), ridx32.size() * 4},
      {wd.data(), wd.size() * 8}};
  size_t off[8] = {0};
  for (int i = 0; i < 7; ++i) off[i + 1] = off[i] + ((std::max<size_t>(parts[i].second, 1) + 255) & ~(size_t)255);
  if (ctx->ev_up) DRPT_HIP(hipEventSynchronize(ctx->ev_up));
  else DRPT_HIP(hipEventCreateWithFlags(&ctx->ev_up, hipEventDisableTiming));
  DRPT_CHECK(pinned_words(ctx, ctx->pin_up, (int)(off[7] / 4)));
  unsigned char* stage = (unsigned char*)ctx->pin_up.host;
  for (int i = 0; i < 7; ++i)
    if (parts[i].second) std::memcpy(stage + off[i], parts[i].first, parts[i].second);
  unsigned char* dev;
  DRPT_CHECK(ws(ctx, "proj.terms", off[7], &dev));
  DRPT_HIP(hipMemcpyAsync(dev, stage, off[7], hipMemcpyHostToDevice, ctx->stream));
  DRPT_HIP(hipEventRecord(ctx->ev_up, ctx->stream));
  const int32_t* dcp = (const int32_t*)(dev + off[0]);
  const uint16_t* dri = (const uint16_t*)(dev + off[1]);
  const float* dw = (const float*)(dev + off[2]);
  const int32_t* dgk = (const int32_t*)(dev + off[3]);
  const int32_t* dgcol = (const int32_t*)(dev + off[4]);
  const int32_t* dri32 = (const int32_t*)(dev + off[5]);
  const double* dwd = (const double*)(dev + off[6]);
  return launch_project(ctx, X, dtype, n, d, dcp, dri, dw, dgk, dgcol, dri32, dwd, L, colptr[L], P,
                        colmax);
}

int drpt_grow_trees(drpt_ctx* ctx, const float* P, const uint32_t* colmax, int64_t n, int ntrees,
                    int depth, int tree_mode, int leaf_max, uint32_t* leaf_heap) {
  if (!ctx || !P || !colmax || !leaf_heap) {
    set_error("grow_trees: null pointer");
    return DRPT_ERR_ARG;
  }
  if (depth <= 0 || (tree_mode == DRPT_TREE_FIXED && depth > std::log2((double)n))) {
    set_error("depth should be in range [1,....,log2(rows)]");
    return DRPT_ERR_DEPTH;
  }
  DRPT_HIP(hipSetDevice(ctx->device));
  return grow_trees_impl(ctx, P, colmax, n, ntrees, depth, tree_mode, leaf_max, leaf_heap);
}

// Final leaves as work items: the points of every leaf contiguous in G->perm
// (tree-major) and items of plan.rpi[bucket] rows; items of trees < first_trees
// come first. Leaf-max builds do this on the device; the reference local stage
// derives its position tables on the host.
struct Grouping {
  const int32_t* perm = nullptr;
  LeafItem* items = nullptr;
  int64_t* range = nullptr;  // device {0, first, first, all, 0, all}
  const int64_t* base = nullptr;  // device [T + 1]: first work item of each tree
  int64_t item_cap = 0;           // capacity of items
  unsigned long long* stats = nullptr;
  double host_gram = -1.0;
  int64_t host_leaves = 0, host_max = 0;
};

static int group_impl(drpt_ctx* ctx, const drpt_params* p, const void* X,
                      const uint32_t* leaf_heap, const ItemPlan& plan, int first_trees,
                      Grouping* G, bool items_only = false) {
  const int T = p->ntrees, D = p->depth, K = p->nn;
  const int64_t n = p->n;
  const int64_t HS = (int64_t)1 << D;
  hipStream_t s = ctx->stream;
  // leaf sizes by heap index
  uint32_t *hcount, *cursor;
  int64_t *hoff, *base, *range;
  int32_t* perm;
  LeafItem* items;
  unsigned long long* stats;
  const int64_t item_cap = (int64_t)T * (n / 64 + std::min<int64_t>(n, HS) + 1);
  DRPT_CHECK(ws(ctx, "group.hcount", T * HS, &hcount));
  DRPT_CHECK(ws(ctx, "group.cursor", T * HS, &cursor));
  DRPT_CHECK(ws(ctx, "group.hoff", T * HS, &hoff));
  DRPT_CHECK(ws(ctx, "group.base", T + 1, &base));
  DRPT_CHECK(ws(ctx, "group.range", 6, &range));
  DRPT_CHECK(ws(ctx, "group.stats", 4, &stats));
  DRPT_CHECK(ws(ctx, "group.perm", T * n, &perm));
  DRPT_CHECK(ws(ctx, "knn.items", item_cap, &items));
  DRPT_CHECK(fill_u32(ctx, (uint32_t*)stats, 0u, 8));
  // items_only (leaf-max builds): new work items over the leaves and point
  // order of the previous call (the scatter's order within a leaf is not
  // reproducible, and leaf-ordered copies made since depend on it)
  if (!items_only || p->local_mode != DRPT_LOCAL_LEAFMAX)
    DRPT_CHECK(launch_count_leaves(ctx, leaf_heap, n, T, D, hcount));
  const int32_t* perm_final = perm;
  double host_gram = -1.0;
  int64_t host_leaves = 0, host_max = 0;

  if (p->local_mode == DRPT_LOCAL_LEAFMAX) {
    // everything on the device: offsets, work items, scatter
    DRPT_CHECK(launch_device_items(ctx, hcount, n, T, D, plan, first_trees, hoff, base, range,
                                   items, stats));
    if (!items_only) DRPT_CHECK(launch_scatter_leaves(ctx, leaf_heap, n, T, D, hoff, cursor, perm));
  } else {
    // reference local stage: the host derives the position tables, so the
    // global leaf sizes come back once
    std::vector<uint32_t> hc((size_t)(T * HS));
    DRPT_HIP(hipMemcpyAsync(hc.data(), hcount, hc.size() * 4, hipMemcpyDeviceToHost, s));
    DRPT_HIP(hipStreamSynchronize(s));
    std::vector<int64_t> hoffh((size_t)(T * HS));
    std::vector<Leaf> leaves;
    for (int t = 0; t < T; ++t) {
      int64_t acc = 0;
      for (int64_t h = 0; h < HS; ++h) {
        const uint32_t c = hc[(size_t)(t * HS + h)];
        hoffh[(size_t)(t * HS + h)] = acc;
        if (c) leaves.push_back({t, (int32_t)c, acc});
        acc += c;
      }
    }
    DRPT_HIP(hipMemcpyAsync(hoff, hoffh.data(), hoffh.size() * 8, hipMemcpyHostToDevice, s));
    DRPT_CHECK(launch_scatter_leaves(ctx, leaf_heap, n, T, D, hoff, cursor, perm));
    // MDRPT::get_global_minimum_leaf_size (mdrpt.cpp:293-319), grow_local_trees (:321-400)
    int64_t gmin = INT64_MAX;
    int max_seg = 0;
    for (auto& l : leaves) {
      gmin = std::min<int64_t>(gmin, l.size);
      max_seg = std::max(max_seg, l.size);
    }
    const int ld = host_local_depth(gmin, K, p->local_offset);
    if (ld <= 0) {
      set_error("local tree depth " + std::to_string(ld) + " (smallest global leaf " +
                std::to_string(gmin) + "): depth should be in range [1,....,log2(rows)]");
      return DRPT_ERR_DEPTH;  // drpt_local.cpp:84-87
    }
    // the local matrices (seed S+1+t, the int-truncated density,
    // mdrpt.cpp:348-352): all zero for a density below 1 (the libstdc++ tie
    // permutation), non-zero for -density >= 1 (the real-key local tree)
    const float dint = (float)(int)p->density;
    const int LL = T * ld;
    std::vector<float> C((size_t)p->d * ld), Call((size_t)p->d * LL);
    int nreal = 0;
    for (int t = 0; t < T; ++t) {
      host_build_projection(p->d, ld, dint, p->seed + 1u + (uint32_t)t, C.data());
      bool real = false;
      for (float c : C) real = real || c != 0.0f;
      nreal += real;
      for (int j = 0; j < p->d; ++j)
        for (int l = 0; l < ld; ++l) Call[(size_t)j * LL + t * ld + l] = C[(size_t)j * ld + l];
    }
    if (nreal != 0 && nreal != T) {
      set_error("local stage: some trees have an all-zero local projection and others not");
      return DRPT_ERR_UNSUPPORTED;
    }
    const bool real_local = nreal == T;
    std::vector<int64_t> seg_start;
    std::vector<int32_t> seg_size, seg_tree;
    std::vector<int64_t> seg_tbl;
    std::vector<int32_t> tables;
    std::map<int, int64_t> tbl_of;
    std::vector<int32_t> sizes((size_t)1 << ld);
    std::vector<Leaf> final_leaves;
    for (auto& l : leaves) {
      seg_start.push_back(l.start);
      seg_size.push_back(l.size);
      seg_tree.push_back(l.tree);
      if (!real_local) {
        auto it = tbl_of.find(l.size);
        if (it == tbl_of.end()) {
          const int64_t off = (int64_t)tables.size();
          tables.resize(tables.size() + l.size);
          host_local_positions(l.size, ld, tables.data() + off);
          it = tbl_of.emplace(l.size, off).first;
        }
        seg_tbl.push_back(it->second);
      }
      host_local_sizes(l.size, ld, sizes.data());
      int64_t acc = l.start;
      for (int32_t sz : sizes) {
        if (sz > 0) final_leaves.push_back({l.tree, sz, acc});
        acc += sz;
      }
    }
    // level-by-level segments of the real-key local tree: at depth l each
    // global leaf holds 2^l nodes with the ceil/floor halving sizes
    std::vector<int64_t> lv_start, lv_ofs(1, 0);
    std::vector<int32_t> lv_size, lv_tree;
    if (real_local) {
      std::vector<int32_t> hs((size_t)1 << ld);
      for (int l = 0; l < ld; ++l) {
        for (auto& lf : leaves) {
          host_local_sizes(lf.size, l, hs.data());
          int64_t acc = lf.start;
          for (int k = 0; k < (1 << l); ++k) {
            if (hs[k] > 1) {
              lv_start.push_back(acc);
              lv_size.push_back(hs[k]);
              lv_tree.push_back(lf.tree);
            }
            acc += hs[k];
          }
        }
        lv_ofs.push_back((int64_t)lv_start.size());
      }
    }
    int64_t* dss;
    int32_t *dsz, *dst;
    DRPT_CHECK(upload(ctx, "local.start", seg_start, &dss));
    DRPT_CHECK(upload(ctx, "local.size", seg_size, &dsz));
    DRPT_CHECK(upload(ctx, "local.tree", seg_tree, &dst));
    if (!real_local) {
      int32_t* dtb;
      int64_t* dtbo;
      DRPT_CHECK(upload(ctx, "local.tblofs", seg_tbl, &dtbo));
      DRPT_CHECK(upload(ctx, "local.tables", tables, &dtb));
      DRPT_CHECK(launch_sort_segments(ctx, perm, n, dss, dsz, dst, (int)leaves.size(), max_seg));
      int32_t* perm2;
      DRPT_CHECK(ws(ctx, "group.perm2", T * n, &perm2));
      DRPT_CHECK(
          launch_apply_local(ctx, perm, perm2, n, dss, dst, dtbo, dsz, (int)leaves.size(), dtb));
      perm_final = perm2;
    } else {
      // local projections of every point, LP[t * ld + l][i]
      // (MDRPT::grow_local_trees' multiply_mat per leaf, mdrpt.cpp:365-369)
      float* LP;
      uint32_t* lcm;
      DRPT_CHECK(ws(ctx, "local.P", (size_t)LL * n, &LP));
      DRPT_CHECK(ws(ctx, "local.colmax", LL, &lcm));
      DRPT_CHECK(drpt_project(ctx, X, p->dtype, n, p->d, Call.data(), LL, LP, lcm));
      std::vector<uint32_t> cmh(LL);
      DRPT_HIP(hipMemcpyAsync(cmh.data(), lcm, 4 * (size_t)LL, hipMemcpyDeviceToHost, s));
      DRPT_HIP(hipStreamSynchronize(s));
      for (uint32_t b : cmh)
        if (b >= 0x7f800000u) {
          set_error("non-finite local projected values");
          return DRPT_ERR_NONFINITE;
        }
      int64_t* dls;
      int32_t *dlz, *dlt;
      DRPT_CHECK(upload(ctx, "local.lv_start", lv_start, &dls));
      DRPT_CHECK(upload(ctx, "local.lv_size", lv_size, &dlz));
      DRPT_CHECK(upload(ctx, "local.lv_tree", lv_tree, &dlt));
      for (int l = 0; l < ld; ++l)
        DRPT_CHECK(launch_sort_keyed(ctx, perm, n, LP, ld, l, dls + lv_ofs[l], dlz + lv_ofs[l],
                                     dlt + lv_ofs[l], (int)(lv_ofs[l + 1] - lv_ofs[l]), max_seg));
    }
    // work items in tree order (tree 0 first), largest leaves first inside a tree
    std::stable_sort(final_leaves.begin(), final_leaves.end(), [](const Leaf& a, const Leaf& b) {
      if (a.tree != b.tree) return a.tree < b.tree;
      return a.size > b.size;
    });
    std::vector<LeafItem> hitems;
    int64_t nfirst = 0;
    host_gram = 0.0;
    for (auto& l : final_leaves) {
      host_gram += 2.0 * p->d * (double)l.size * (double)l.size;
      host_max = std::max<int64_t>(host_max, l.size);
      const int rpi = plan.rpi[size_bucket(l.size)];
      for (int64_t r0 = 0; r0 < l.size; r0 += rpi) {
        hitems.push_back({l.tree, l.size, l.start, (int32_t)r0,
                          (int32_t)std::min<int64_t>(rpi, l.size - r0)});
        if (l.tree < first_trees) ++nfirst;
      }
    }
    host_leaves = (int64_t)final_leaves.size();
    if ((int64_t)hitems.size() > item_cap) {
      set_error("internal: work item capacity exceeded");
      return DRPT_ERR_ARG;
    }
    const int64_t ni = (int64_t)hitems.size();
    std::vector<int64_t> hr = {0, nfirst, nfirst, ni, 0, ni};
    std::vector<int64_t> hbase((size_t)T + 1, 0);
    for (auto& it : hitems) ++hbase[(size_t)it.tree + 1];
    for (int t = 0; t < T; ++t) hbase[(size_t)t + 1] += hbase[(size_t)t];
    DRPT_HIP(hipMemcpyAsync(base, hbase.data(), hbase.size() * 8, hipMemcpyHostToDevice, s));
    DRPT_HIP(hipMemcpyAsync(items, hitems.data(), hitems.size() * sizeof(LeafItem),
                            hipMemcpyHostToDevice, s));
    DRPT_HIP(hipMemcpyAsync(range, hr.data(), 6 * 8, hipMemcpyHostToDevice, s));
    DRPT_HIP(hipStreamSynchronize(s));  // host vectors go out of scope
  }

  G->perm = perm_final;
  G->items = items;
  G->item_cap = item_cap;
  G->range = range;
  G->base = base;
  G->stats = stats;
  G->host_gram = host_gram;
  G->host_leaves = host_leaves;
  G->host_max = host_max;
  return DRPT_OK;
}

// leaf statistics of a grouping into timings, from the device counters st
// (copied back by the caller) or the host's own figures
static void leaf_stats_of(const drpt_params* p, const Grouping& G, const unsigned long long* st,
                          drpt_timings* tm) {
  if (!tm) return;
  if (G.host_gram >= 0.0) {
    tm->gram_flops = G.host_gram;
    tm->n_leaves = G.host_leaves;
    tm->max_leaf = G.host_max;
  } else {
    tm->gram_flops = 2.0 * p->d * (double)st[0];
    tm->n_leaves = (int64_t)st[2];
    tm->max_leaf = (int64_t)st[1];
  }
}

// leaf statistics of a grouping into timings (reads the device counters: syncs)
static int leaf_stats(drpt_ctx* ctx, const drpt_params* p, const Grouping& G, drpt_timings* tm) {
  unsigned long long st[4] = {0, 0, 0, 0};
  DRPT_HIP(hipMemcpyAsync(st, G.stats, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
  DRPT_HIP(hipStreamSynchronize(ctx->stream));
  leaf_stats_of(p, G, st, tm);
  return DRPT_OK;
}

// quantised copy (bf16, or fp8 e4m3 with a power-of-two scale per row),
// |x~|^2 and |x~ - x| per point, then their leaf-ordered copies
static int prep_lowp(drpt_ctx* ctx, const drpt_params* p, const void* X, const int32_t* perm,
                     KnnArgs* a, uint8_t* xi8 = nullptr, int32_t* notu8 = nullptr) {
  const int64_t n = p->n;
  const bool fp8 = p->precision == DRPT_FP8;
  const std::string tag = fp8 ? "lowp8." : "lowp.";  // both may be live in one build
  uint8_t* xq;
  float2* nbe;
  float *lnbT, *lerrT;
  int8_t *pexp = nullptr, *lexpT = nullptr;
  DRPT_CHECK(ws(ctx, (tag + "x").c_str(), (size_t)n * p->d * (fp8 ? 1 : 2), &xq));
  DRPT_CHECK(ws(ctx, (tag + "nbe").c_str(), n, &nbe));
  if (fp8) DRPT_CHECK(ws(ctx, (tag + "pexp").c_str(), n, &pexp));
  DRPT_CHECK(launch_prep_lowp(ctx, X, p->dtype, n, p->d, p->precision, xq, nbe, pexp, xi8, notu8));
  DRPT_CHECK(ws(ctx, (tag + "lnbT").c_str(), (size_t)p->ntrees * n, &lnbT));
  DRPT_CHECK(ws(ctx, (tag + "lerrT").c_str(), (size_t)p->ntrees * n, &lerrT));
  if (fp8) DRPT_CHECK(ws(ctx, (tag + "lexpT").c_str(), (size_t)p->ntrees * n, &lexpT));
  // with the int8 instances possible, the copies only when the data are not
  // integers in [0, 255] (the int8 instances read none; the streaming kernel
  // then reads nbe)
  const int32_t* gate = notu8 != nullptr && ctx->use_int8 && pool_leaf_int8(p->d) ? notu8 : nullptr;
  DRPT_CHECK(launch_leaf_gather(ctx, perm, (int64_t)p->ntrees * n, nbe, pexp, lnbT, lerrT, lexpT,
                                gate));
  a->X = X;
  a->Xq = xq;
  a->precision = p->precision;
  a->lexpT = lexpT;
  a->lnbT = lnbT;
  a->lerrT = lerrT;
  a->nbe = nbe;
  a->lnb_gate = gate;
  a->perm = perm;
  a->n = n;
  a->d = p->d;
  a->K = p->nn;
  return DRPT_OK;
}

static int tree_knn_impl(drpt_ctx* ctx, const drpt_params* p, const void* X,
                         const uint32_t* leaf_heap, int32_t* cand_idx, float* cand_dist,
                         drpt_timings* timings);

int drpt_tree_knn(drpt_ctx* ctx, const drpt_params* p, const void* X, const uint32_t* leaf_heap,
                  int32_t* cand_idx, float* cand_dist, drpt_timings* timings) {
  return tree_knn_impl(ctx, p, X, leaf_heap, cand_idx, cand_dist, timings);
}

// Exact per-tree lists (FP32: every pair; BF16: bound pass + filter pass).
static int tree_knn_impl(drpt_ctx* ctx, const drpt_params* p, const void* X,
                         const uint32_t* leaf_heap, int32_t* cand_idx, float* cand_dist,
                         drpt_timings* timings) {
  if (!ctx || !X || !leaf_heap || !cand_idx || !cand_dist) {
    set_error("tree_knn: null pointer");
    return DRPT_ERR_ARG;
  }
  DRPT_CHECK(validate(p));
  DRPT_HIP(hipSetDevice(ctx->device));
  const int T = p->ntrees;
  const int64_t n = p->n;
  hipStream_t s = ctx->stream;
  DRPT_HIP(hipEventRecord(ctx->ev[6], s));
  Grouping G;
  DRPT_CHECK(group_impl(ctx, p, X, leaf_heap, uniform_plan(64), T, &G));
  drpt_params pp = *p;
  pp.kprime = kprime_of(p);
  int32_t *fb = nullptr, *nfb = nullptr;
  const int64_t fb_cap = std::min<int64_t>((int64_t)T * n, (int64_t)1 << 22);
  if (pp.precision == DRPT_FP32) {
    DRPT_HIP(hipEventRecord(ctx->ev[7], s));
    DRPT_HIP(hipEventRecord(ctx->ev[9], s));
    DRPT_CHECK(launch_leaf_knn(ctx, &pp, X, G.perm, G.items, G.range + 4, cand_idx, cand_dist));
    DRPT_HIP(hipEventRecord(ctx->ev[8], s));
  } else {
    if (pp.kprime < p->nn) {
      set_error("kprime must be >= nn");
      return DRPT_ERR_ARG;
    }
    KnnArgs a{};
    DRPT_CHECK(prep_lowp(ctx, &pp, X, G.perm, &a));
    DRPT_CHECK(ws(ctx, "knn.fallback", 2 * fb_cap, &fb));
    DRPT_CHECK(ws(ctx, "knn.nfallback", 1, &nfb));
    DRPT_CHECK(fill_u32(ctx, (uint32_t*)nfb, 0u, 1));
    a.items = G.items;
    a.range = G.range + 4;
    a.size_lo = 0;
    a.size_hi = INT_MAX;
    a.ring = 0;
    a.out_idx = cand_idx;
    a.out_dist = cand_dist;
    a.fallback = fb;
    a.nfallback = nfb;
    a.fb_cap = fb_cap;
    DRPT_HIP(hipEventRecord(ctx->ev[7], s));
    DRPT_HIP(hipEventRecord(ctx->ev[9], s));
    DRPT_CHECK(launch_lowp_knn(ctx, &pp, kModeTree, a));
    DRPT_HIP(hipEventRecord(ctx->ev[8], s));
    // rows whose candidate list overflowed, or of leaves too large to stage
    DRPT_CHECK(launch_fallback(ctx, &pp, X, G.perm, G.items, fb, nfb, fb_cap, cand_idx, cand_dist));
  }
  DRPT_HIP(hipEventRecord(ctx->ev[5], s));
  int32_t nfallback = 0;
  if (nfb) DRPT_HIP(hipMemcpyAsync(&nfallback, nfb, 4, hipMemcpyDeviceToHost, s));
  DRPT_CHECK(leaf_stats(ctx, p, G, timings));
  if (nfallback > fb_cap) {
    set_error("too many candidate-list overflows for the fallback list");
    return DRPT_ERR_UNSUPPORTED;
  }
  if (timings) {
    timings->group_ms = ms_between(ctx->ev[6], ctx->ev[7]);
    timings->knn_ms = ms_between(ctx->ev[7], ctx->ev[5]);
    timings->knn_kernel_ms = ms_between(ctx->ev[9], ctx->ev[8]);
    timings->n_fallback_rows = nfallback;
  }
  return DRPT_OK;
}

// Trees whose merged widened candidates bound the K-th distance of every row.
constexpr int kFirstTrees = 4;
// Pool slots per row (more survivors send the row to the exact row path): the
// pool pass drops a row's known ids when K <= 16 (ids staged in LDS); with
// larger K they stay in the pool and the row merge drops them, so the pool
// is larger (a row merge of up to 1024 entries is still far cheaper than the
// exact path over every leafmate of every tree).
inline int pool_cap_of(int K) { return K <= 16 ? 256 : 1024; }

// Low-precision graph (build_graph): the final lists straight from the leaves.
//  1. kModeCand on trees < a: k' widened candidates per row and tree (with
//     their approximate distances)
//  2. row merge: distinct candidates, exact D of the 2K best approximate ones,
//     K best -> known list; its K-th distance B(i) bounds the K-th distance of
//     the final list (any K exact distances do)
//  3. kModePool on every tree: leafmates with LB <= B(i) not in the known
//     list go to the row's pool (every member of the final top K passes)
//  4. row merge of pools and known list -> final lists; rows whose pool
//     overflowed are redone exactly over every tree
// Trees per pooled filter pass: a row's pool holds the survivors of every
// tree of the group (a neighbour repeats once per tree whose leaf holds it),
// so larger K takes smaller groups; one group when K <= 10.
static int tree_group_auto(int T, int K) {
  return std::max(1, std::min(T, std::max(4, 160 / std::max(K, 1))));
}

// The low-precision pipeline's state: grouping, quantised inputs, buffers.
// The candidate pass runs at the build's precision (pc, ac); the pooled
// filter passes that certify the lists run in BF16 (pp, a): an FP8 error
// radius (about 3% of |x| per row) admits so many leafmates on clustered data
// that most rows would leave their pools.
struct LowpRun {
  drpt_params pp, pc;
  Grouping G;
  KnnArgs a{}, ac{};
  int T = 0, K = 0, KP = 0, first = 0, pool_cap = 0;
  int64_t n = 0;
  uint64_t* s1 = nullptr;
  int32_t *ki = nullptr, *pool = nullptr, *pcnt = nullptr, *oflag = nullptr, *olist = nullptr,
          *novf = nullptr;
  float *kd = nullptr, *boundT = nullptr;
  float* bnd = nullptr;  // [n] K-th distance of the current known lists (the bound)
  // rows the merges read: X itself, or (locality) X in tree 0's leaf order
  // with xpos[i] the position of point i in it
  const void* Xm = nullptr;
  const int32_t* xpos = nullptr;
  int32_t* lossy = nullptr;  // device flag: some row differs from its bf16 copy
  // float X: its int8 copy (x - 128) and a device flag, some value not an
  // integer in [0, 255]
  uint8_t* xi8 = nullptr;
  int32_t* notu8 = nullptr;
};

static std::string plan_key(const drpt_params* p) {
  char k[256];
  snprintf(k, sizeof k, "%lld %d %d %d %d %d %d %d %d %d %d", (long long)p->n, p->d, p->dtype,
           p->precision, p->nn, p->ntrees, p->depth, p->leaf_max, p->local_mode, kprime_of(p),
           p->tree_group);
  return k;
}

// whole leaves; the pooled filter pass with each leaf resident in LDS where it
// fits, else (and for the candidate pass) the LDS tile ring
static ItemPlan lowp_default_plan() {
  ItemPlan p = uniform_plan(1 << 30);
  for (int b = 0; b < kSizeBuckets; ++b) p.ring[b] = 2;
  return p;
}

static ItemPlan lowp_plan_of(drpt_ctx* ctx, const drpt_params* p) {
  auto it = ctx->plans.find(plan_key(p));
  if (it == ctx->plans.end()) return lowp_default_plan();
  ItemPlan plan;
  for (int b = 0; b < kSizeBuckets; ++b) {
    plan.rpi[b] = it->second[b];
    plan.ring[b] = it->second[kSizeBuckets + b];
  }
  plan.tree_group = it->second[2 * kSizeBuckets];
  plan.first_trees = it->second[2 * kSizeBuckets + 1];
  plan.rerank = it->second[2 * kSizeBuckets + 2];
  return plan;
}

// grouping with `plan`, quantised copies, buffers; the pool and overflow
// state cleared
static int lowp_setup(drpt_ctx* ctx, const drpt_params* p, const void* X,
                      const uint32_t* leaf_heap, const ItemPlan& plan, LowpRun* r) {
  r->pp = *p;
  r->pp.kprime = kprime_of(p);
  r->pc = r->pp;
  r->pp.precision = DRPT_BF16;
  r->T = p->ntrees;
  r->K = p->nn;
  r->n = p->n;
  if (r->pp.kprime < r->K) {
    set_error("kprime must be >= nn");
    return DRPT_ERR_ARG;
  }
  r->KP = lowp_slots(&r->pp);
  r->first = std::max(1, std::min(std::min(r->T, plan.first_trees > 0 ? plan.first_trees
                                                                      : kFirstTrees),
                                  256 / r->KP));
  // the LDS-resident pool pass leaves known ids to the row merge: larger pools
  r->pool_cap = pool_leaf_rows(ctx, p->d, kModePool) > 0 ? std::max(512, pool_cap_of(r->K))
                                                         : pool_cap_of(r->K);
  const int64_t n = r->n;
  DRPT_CHECK(group_impl(ctx, p, X, leaf_heap, plan, r->first, &r->G));
  r->Xm = X;
  r->xpos = nullptr;
  if (p->locality) {
    const size_t rowbytes = (size_t)p->d * (p->dtype == DRPT_U8 ? 1 : 4);
    void* Xp;
    int32_t* pos;
    DRPT_CHECK(ws_get(ctx, "loc.X", (size_t)n * rowbytes + 16, &Xp));
    DRPT_CHECK(ws(ctx, "loc.pos", (size_t)n, &pos));
    DRPT_CHECK(launch_permute_rows(ctx, X, rowbytes, r->G.perm, n, Xp, pos));
    r->Xm = Xp;
    r->xpos = pos;
  }
  DRPT_HIP(hipEventRecord(ctx->evp[0], ctx->stream));
  if (p->dtype == DRPT_F32) {
    DRPT_CHECK(ws(ctx, "lowp.xi8", (size_t)n * p->d, &r->xi8));
    DRPT_CHECK(ws(ctx, "lowp.notu8", 1, &r->notu8));
    DRPT_CHECK(fill_u32(ctx, (uint32_t*)r->notu8, 0u, 1));
  }
  DRPT_CHECK(prep_lowp(ctx, &r->pp, X, r->G.perm, &r->a, r->xi8, r->notu8));
  if (ctx->use_int8) {  // the LDS-resident passes' int8 instances
    r->a.Xi8 = r->xi8;
    r->a.notu8 = r->notu8;
  }
  {
    float2* nbe;
    DRPT_CHECK(ws(ctx, "lowp.nbe", (size_t)n, &nbe));
    DRPT_CHECK(ws(ctx, "lowp.lossy", 1, &r->lossy));
    DRPT_CHECK(fill_u32(ctx, (uint32_t*)r->lossy, 0u, 1));
    DRPT_CHECK(launch_lossy_flag(ctx, nbe, n, r->lossy));
  }
  if (r->pc.precision != r->pp.precision) {
    DRPT_CHECK(prep_lowp(ctx, &r->pc, X, r->G.perm, &r->ac));
  } else {
    r->ac = r->a;
  }
  r->ac.items = r->G.items;
  DRPT_CHECK(ws(ctx, "knn.s1", (size_t)r->first * n * r->KP, &r->s1));
  DRPT_CHECK(ws(ctx, "knn.known_idx", (size_t)n * r->K, &r->ki));
  DRPT_CHECK(ws(ctx, "knn.known_dist", (size_t)n * r->K, &r->kd));
  DRPT_CHECK(ws(ctx, "knn.boundT", (size_t)r->T * n, &r->boundT));
  DRPT_CHECK(ws(ctx, "knn.bound", (size_t)n, &r->bnd));
  DRPT_CHECK(ws(ctx, "knn.pool", (size_t)n * r->pool_cap, &r->pool));
  DRPT_CHECK(ws(ctx, "knn.pool_cnt", n, &r->pcnt));
  DRPT_CHECK(ws(ctx, "knn.ovf_flag", n, &r->oflag));
  DRPT_CHECK(ws(ctx, "knn.ovf_list", n, &r->olist));
  DRPT_CHECK(ws(ctx, "knn.novf", 1, &r->novf));
  DRPT_CHECK(fill3_u32(ctx, (uint32_t*)r->pcnt, 0u, n, (uint32_t*)r->oflag, 0u, n,
                       (uint32_t*)r->novf, 0u, 1));
  KnnArgs& a = r->a;
  a.items = r->G.items;
  a.out_idx = r->pool;
  a.out_dist = nullptr;
  a.boundT = r->boundT;
  a.pool_cnt = r->pcnt;
  a.pool_cap = r->pool_cap;
  a.ovf_flag = r->oflag;
  a.ovf_list = r->olist;
  a.novf = r->novf;
  return DRPT_OK;
}

// one low-precision pass over items [range[0], range[1]): a single launch when
// every size bucket uses the same Gram pipeline, else one launch per bucket
// (items of other buckets skipped); bucket >= 0 restricts it to that bucket
static int lowp_pass(drpt_ctx* ctx, LowpRun* r, int mode, const ItemPlan& plan,
                     const int64_t* range, int bucket = -1) {
  KnnArgs a = mode == kModeCand ? r->ac : r->a;
  const drpt_params* pr = mode == kModeCand ? &r->pc : &r->pp;
  a.items = r->G.items;
  a.range = range;
  a.item_cap = r->G.item_cap;
  // leaves of sizes [lo, hi] with Gram pipeline `ring`: the LDS-resident pool
  // pass (ring 2) takes those it can stage, the streaming kernel the rest
  // (candidate pass: BF16 rows and k' in {12, 16, 24} only)
  const int kp = lowp_slots(pr);
  const bool pl_ok = mode == kModePool ||
                     (mode == kModeCand && pr->precision == DRPT_BF16 && (kp == 12 || kp == 16 || kp == 24));
  const int pl_rows = pl_ok ? pool_leaf_rows(ctx, pr->d, mode) : 0;
  auto run = [&](int lo, int hi, int ring) -> int {
    KnnArgs b = a;
    if (ring == 2 && pl_rows > 0 && lo <= pl_rows) {
      b.size_lo = lo;
      b.size_hi = std::min(hi, pl_rows);
      DRPT_CHECK(launch_pool_leaf(ctx, mode, kp, b));
      if (hi <= pl_rows) return DRPT_OK;
      lo = pl_rows + 1;
      // the few larger leaves: their items gathered first, so that the
      // streaming kernel does not walk (and skip) every other item
      LeafItem* big;
      int64_t* big_range;
      DRPT_CHECK(ws(ctx, "knn.items_big", (size_t)r->G.item_cap, &big));
      DRPT_CHECK(ws(ctx, "knn.range_big", 2, &big_range));
      DRPT_CHECK(launch_select_items(ctx, a.items, a.range, lo, hi, r->G.item_cap, big, big_range));
      b.items = big;
      b.range = big_range;
    }
    b.size_lo = lo;
    b.size_hi = hi;
    b.ring = ring == 0 ? 0 : 1;
    return launch_lowp_knn(ctx, pr, mode, b);
  };
  bool uniform = true;
  for (int b = 1; b < kSizeBuckets; ++b) uniform = uniform && plan.ring[b] == plan.ring[0];
  if (bucket < 0 && uniform) return run(0, INT_MAX, plan.ring[0]);
  for (int b = 0; b < kSizeBuckets; ++b) {
    if (bucket >= 0 && b != bucket) continue;
    int lo, hi;
    bucket_sizes(b, &lo, &hi);
    DRPT_CHECK(run(lo, hi, plan.ring[b]));
  }
  return DRPT_OK;
}

// 1. widened candidates of the first trees; 2. known lists (ki, kd) and bounds
static int lowp_known(drpt_ctx* ctx, LowpRun* r, const void* X, const ItemPlan& plan) {
  r->ac.out_key = r->s1;
  r->ac.tree0 = 0;
  DRPT_CHECK(lowp_pass(ctx, r, kModeCand, plan, r->G.range + 0));
  DRPT_HIP(hipEventRecord(ctx->evp[1], ctx->stream));
  if (plan.rerank < 0) {
    // bounds from the candidates' approximate distances (no exact distance):
    // the known lists start empty and the first pool merge evaluates every
    // distinct pool entry
    float2* nbe;
    DRPT_CHECK(ws(ctx, r->pc.precision == DRPT_FP8 ? "lowp8.nbe" : "lowp.nbe", (size_t)r->n, &nbe));
    DRPT_CHECK(launch_row_bound(ctx, r->s1, r->n * r->KP, r->KP, r->first, r->n, r->pp.d, r->K,
                                nbe, r->bnd));
    // ids -1, distances +inf
    DRPT_CHECK(fill3_u32(ctx, (uint32_t*)r->ki, 0xFFFFFFFFu, (size_t)r->n * r->K,
                         (uint32_t*)r->kd, 0x7F800000u, (size_t)r->n * r->K));
    DRPT_HIP(hipEventRecord(ctx->evp[2], ctx->stream));
    return DRPT_OK;
  }
  RowMergeArgs m{};
  m.X = r->Xm;
  m.xpos = r->xpos;
  m.n = r->n;
  m.d = r->pp.d;
  m.K = r->K;
  m.src_key = r->s1;
  m.src_tree_stride = r->n * r->KP;
  m.S = r->KP;
  m.ntsrc = r->first;
  // at least K exact distances per row: the K-th of them bounds the final K-th
  const int rr = plan.rerank > 0 ? std::max(r->K, std::min(plan.rerank, 2 * r->K)) : 2 * r->K;
  m.rerank = std::min(rr, r->first * r->KP);
  m.out_idx = r->ki;
  m.out_dist = r->kd;
  m.out_bound = r->bnd;
  m.order = r->G.perm;  // tree 0's leaf order
  DRPT_CHECK(launch_row_merge(ctx, r->pp.dtype, m));
  DRPT_HIP(hipEventRecord(ctx->evp[2], ctx->stream));
  return DRPT_OK;
}

static int graph_knn_lowp(drpt_ctx* ctx, const drpt_params* p, const void* X,
                          const uint32_t* leaf_heap, int32_t* out_idx, float* out_dist,
                          drpt_timings* timings) {
  hipStream_t s = ctx->stream;
  const ItemPlan plan = lowp_plan_of(ctx, p);
  DRPT_HIP(hipEventRecord(ctx->ev[6], s));
  LowpRun r;
  DRPT_CHECK(lowp_setup(ctx, p, X, leaf_heap, plan, &r));
  const int T = r.T, K = r.K;
  const int64_t n = r.n;
  const Grouping& G = r.G;
  DRPT_HIP(hipEventRecord(ctx->ev[7], s));
  DRPT_CHECK(lowp_known(ctx, &r, X, plan));
  // 3. pools over groups of trees, each merged into the known lists before
  // the next group (duplicates across the group's trees removed, the bound
  // tightened); the last merge writes the output
  const int gsz = p->tree_group > 0   ? std::min(p->tree_group, T)
                  : plan.tree_group > 0 ? std::min(plan.tree_group, T)
                                        : tree_group_auto(T, K);
  const int ngroups = (T + gsz - 1) / gsz;
  int64_t* granges;
  int32_t* ki2 = nullptr;
  float* kd2 = nullptr;
  DRPT_CHECK(ws(ctx, "knn.granges", 2 * (size_t)ngroups, &granges));
  if (ngroups > 1) {
    DRPT_CHECK(ws(ctx, "knn.known_idx2", (size_t)n * K, &ki2));
    DRPT_CHECK(ws(ctx, "knn.known_dist2", (size_t)n * K, &kd2));
  }
  int32_t* heavy = nullptr;  // rows whose pools hold more than 256 entries
  int32_t* nheavy = nullptr;
  if (r.pool_cap > 256) {
    DRPT_CHECK(ws(ctx, "knn.heavy", (size_t)n, &heavy));
    DRPT_CHECK(ws(ctx, "knn.nheavy", 1, &nheavy));
  }
  while ((int)ctx->ev_pass.size() < 4 * ngroups) {
    hipEvent_t e;
    DRPT_HIP(hipEventCreate(&e));
    ctx->ev_pass.push_back(e);
  }
  for (int g = 0; g < ngroups; ++g)
    DRPT_CHECK(launch_tree_range(ctx, G.base, g * gsz, std::min(T, (g + 1) * gsz), granges + 2 * g));
  unsigned long long* pool_sum;  // entries the pooled passes appended (a statistic)
  DRPT_CHECK(ws(ctx, "knn.pool_sum", 1, &pool_sum));
  DRPT_CHECK(fill_u32(ctx, (uint32_t*)pool_sum, 0u, 2));
  int32_t* kin = r.ki;  // known lists in, merged lists out (ping-pong)
  float* kdn = r.kd;
  for (int g = 0; g < ngroups; ++g) {
    const int t0 = g * gsz, t1 = std::min(T, t0 + gsz);
    const bool last = g + 1 == ngroups;
    DRPT_HIP(hipEventRecord(ctx->ev_pass[4 * g + 3], s));
    DRPT_CHECK(launch_leaf_gather_val(ctx, G.perm + (int64_t)t0 * n, (int64_t)(t1 - t0) * n, r.bnd,
                                      1, 0, r.boundT + (int64_t)t0 * n));
    if (g > 0) DRPT_CHECK(fill_u32(ctx, (uint32_t*)r.pcnt, 0u, n));
    DRPT_HIP(hipEventRecord(ctx->ev_pass[4 * g], s));
    r.a.known = kin;
    DRPT_CHECK(lowp_pass(ctx, &r, kModePool, plan, granges + 2 * g));
    DRPT_HIP(hipEventRecord(ctx->ev_pass[4 * g + 1], s));
    DRPT_CHECK(sum_u32(ctx, (const uint32_t*)r.pcnt, n, pool_sum));
    RowMergeArgs f{};
    f.X = r.Xm;
    f.xpos = r.xpos;
    f.Xb = static_cast<const uint16_t*>(r.a.Xq);  // the pooled passes' bf16 copy
    f.lossy = r.lossy;
    f.Xi8 = r.xi8;
    f.notu8 = r.notu8;
    f.n = n;
    f.d = p->d;
    f.K = K;
    f.src = r.pool;
    f.src_tree_stride = 0;
    f.S = r.pool_cap;
    f.ntsrc = 1;
    f.cnt_tree = r.pcnt;
    f.known_idx = kin;
    f.known_dist = kdn;
    f.skip_flag = r.oflag;
    f.out_idx = last ? out_idx : (kin == r.ki ? ki2 : r.ki);
    f.out_dist = last ? out_dist : (kdn == r.kd ? kd2 : r.kd);
    f.out_bound = last ? nullptr : r.bnd;
    f.order = G.perm;
    f.heavy = heavy;
    f.nheavy = nheavy;
    DRPT_CHECK(launch_row_merge(ctx, p->dtype, f));
    DRPT_HIP(hipEventRecord(ctx->ev_pass[4 * g + 2], s));
    kin = f.out_idx;
    kdn = f.out_dist;
  }
  // overflow rows, pool entries, leaf statistics and the integer flag: one
  // small kernel's stores into pinned memory, one wait
  DRPT_CHECK(pinned_words(ctx, ctx->pin_st, 12));
  const bool has_u8 = r.a.notu8 != nullptr;
  DRPT_CHECK(words_to_host(ctx, ctx->pin_st.dev, (const uint32_t*)r.novf, 1,
                           (const uint32_t*)pool_sum, 2, (const uint32_t*)G.stats, 8,
                           has_u8 ? (const uint32_t*)r.notu8 : nullptr, has_u8 ? 1 : 0));
  DRPT_HIP(hipStreamSynchronize(s));
  const uint32_t* hw = ctx->pin_st.host;
  const int32_t nov = (int32_t)hw[0];
  const int32_t notu8 = has_u8 ? (int32_t)hw[11] : 1;
  const unsigned long long npool = (unsigned long long)hw[1] | ((unsigned long long)hw[2] << 32);
  unsigned long long gst[4];
  for (int k = 0; k < 4; ++k)
    gst[k] = (unsigned long long)hw[3 + 2 * k] | ((unsigned long long)hw[4 + 2 * k] << 32);
  if (nov > 0)
    DRPT_CHECK(launch_row_exact(ctx, &r.pp, X, G.perm, G.items, G.range + 4, r.olist, r.novf, nov,
                                out_idx, out_dist));
  DRPT_HIP(hipEventRecord(ctx->ev[5], s));
  DRPT_HIP(hipEventSynchronize(ctx->ev[5]));  // the stage times below read it
  leaf_stats_of(p, G, gst, timings);
  drpt_phases& ph = ctx->last_phases;
  ph = drpt_phases{};
  ph.group_ms = ms_between(ctx->ev[6], ctx->evp[0]);
  ph.prep_ms = ms_between(ctx->evp[0], ctx->ev[7]);
  ph.cand_ms = ms_between(ctx->ev[7], ctx->evp[1]);
  ph.known_ms = ms_between(ctx->evp[1], ctx->evp[2]);
  for (int g = 0; g < ngroups; ++g) {
    ph.bound_ms += ms_between(ctx->ev_pass[4 * g + 3], ctx->ev_pass[4 * g]);
    ph.pool_ms += ms_between(ctx->ev_pass[4 * g], ctx->ev_pass[4 * g + 1]);
    ph.pool_merge_ms += ms_between(ctx->ev_pass[4 * g + 1], ctx->ev_pass[4 * g + 2]);
  }
  ph.exact_ms = ms_between(ctx->ev_pass[4 * (ngroups - 1) + 2], ctx->ev[5]);
  ph.tree_groups = ngroups;
  ph.first_trees = r.first;
  ph.pool_cap = r.pool_cap;
  ph.rerank = plan.rerank < 0 ? -1 : plan.rerank > 0 ? std::max(K, std::min(plan.rerank, 2 * K)) : 2 * K;
  ph.pool_entries = (int64_t)npool;
  bool lds_pass = false;
  for (int b = 0; b < kSizeBuckets; ++b) lds_pass = lds_pass || plan.ring[b] == 2;
  ph.int8_passes = notu8 == 0 && lds_pass && pool_leaf_int8(p->d) &&
                   pool_leaf_rows(ctx, p->d, kModePool) > 0;
  if (timings) {
    timings->group_ms = ms_between(ctx->ev[6], ctx->ev[7]);
    timings->knn_ms = ms_between(ctx->ev[7], ctx->ev[5]);
    timings->knn_kernel_ms = ph.pool_ms;
    timings->n_fallback_rows = nov;
  }
  return DRPT_OK;
}

int drpt_merge(drpt_ctx* ctx, int64_t n, int T, int K, const int32_t* cand_idx,
               const float* cand_dist, int32_t* out_idx, float* out_dist) {
  if (!ctx || !cand_idx || !cand_dist || !out_idx || !out_dist || n < 1 || T < 1 || K < 1) {
    set_error("merge: invalid arguments");
    return DRPT_ERR_ARG;
  }
  DRPT_HIP(hipSetDevice(ctx->device));
  return launch_merge(ctx, n, T, K, cand_idx, cand_dist, out_idx, out_dist);
}

// projection and trees of a build: leaf heap index of every point, [T][n]
// (events 0-2 bracket the two stages)
static int build_front(drpt_ctx* ctx, const drpt_params* p, const void* X, uint32_t** heap_out) {
  const int T = p->ntrees, D = p->depth, L = T * D;
  const int64_t n = p->n;
  hipStream_t s = ctx->stream;
  std::vector<float> R((size_t)p->d * L);
  host_build_projection(p->d, L, p->density, p->seed, R.data());
  float* P;
  uint32_t *colmax, *heap;
  DRPT_CHECK(ws(ctx, "build.P", (size_t)L * n, &P));
  DRPT_CHECK(ws(ctx, "build.colmax", L, &colmax));
  DRPT_CHECK(ws(ctx, "build.heap", (size_t)T * n, &heap));
  DRPT_HIP(hipEventRecord(ctx->ev[0], s));
  DRPT_CHECK(drpt_project(ctx, X, p->dtype, n, p->d, R.data(), L, P, colmax));
  DRPT_HIP(hipEventRecord(ctx->ev[1], s));
  // a non-finite projection cannot be split (the fixed-point statistics need
  // finite values): the column maxima come back to pinned memory while the
  // splits run (their work is bounded whatever the values), and the error is
  // returned before the leaf stage
  DRPT_CHECK(pinned_words(ctx, ctx->pin_cm, L));
  if (!ctx->ev_cm) DRPT_HIP(hipEventCreateWithFlags(&ctx->ev_cm, hipEventDisableTiming));
  DRPT_CHECK(words_to_host(ctx, ctx->pin_cm.dev, colmax, L));
  DRPT_HIP(hipEventRecord(ctx->ev_cm, s));
  const int mode = p->local_mode == DRPT_LOCAL_REFERENCE ? DRPT_TREE_FIXED : DRPT_TREE_LEAFMAX;
  DRPT_CHECK(grow_trees_impl(ctx, P, colmax, n, T, D, mode, p->leaf_max, heap));
  DRPT_HIP(hipEventRecord(ctx->ev[2], s));
  DRPT_HIP(hipEventSynchronize(ctx->ev_cm));
  for (int c = 0; c < L; ++c)
    if (ctx->pin_cm.host[c] >= 0x7f800000u) {
      set_error("non-finite projected values (input contains inf/nan or overflows float)");
      return DRPT_ERR_NONFINITE;
    }
  *heap_out = heap;
  return DRPT_OK;
}

int drpt_build_graph(drpt_ctx* ctx, const drpt_params* p, const void* X, int32_t* out_idx,
                     float* out_dist, drpt_timings* timings) {
  if (!ctx || !X || !out_idx || !out_dist) {
    set_error("build_graph: null pointer");
    return DRPT_ERR_ARG;
  }
  DRPT_CHECK(validate(p));
  DRPT_HIP(hipSetDevice(ctx->device));
  const int T = p->ntrees, K = p->nn;
  const int64_t n = p->n;
  hipStream_t s = ctx->stream;
  uint32_t* heap;
  DRPT_CHECK(build_front(ctx, p, X, &heap));
  drpt_timings tk{};
  bool fp32 = false;
  if (p->precision == DRPT_FP32) {
    // per-tree lists (MDRPT::calculate_nns), then the merge (:180-198)
    int32_t* ci;
    float* cd;
    DRPT_CHECK(ws(ctx, "build.cand_idx", (size_t)T * n * K, &ci));
    DRPT_CHECK(ws(ctx, "build.cand_dist", (size_t)T * n * K, &cd));
    DRPT_CHECK(tree_knn_impl(ctx, p, X, heap, ci, cd, &tk));
    DRPT_HIP(hipEventRecord(ctx->ev[3], s));
    DRPT_CHECK(launch_merge(ctx, n, T, K, ci, cd, out_idx, out_dist));
    fp32 = true;
  } else {
    // the merged lists directly (the merge is the last row merge)
    DRPT_CHECK(graph_knn_lowp(ctx, p, X, heap, out_idx, out_dist, &tk));
    DRPT_HIP(hipEventRecord(ctx->ev[3], s));
  }
  DRPT_HIP(hipEventRecord(ctx->ev[4], s));
  DRPT_HIP(hipStreamSynchronize(s));
  if (fp32) {
    drpt_phases& ph = ctx->last_phases;
    ph = drpt_phases{};
    ph.group_ms = tk.group_ms;
    ph.pool_ms = tk.knn_kernel_ms;
    ph.exact_ms = ms_between(ctx->ev[8], ctx->ev[5]);
    ph.pool_merge_ms = ms_between(ctx->ev[3], ctx->ev[4]);
  }
  if (timings) {
    *timings = tk;
    timings->project_ms = ms_between(ctx->ev[0], ctx->ev[1]);
    timings->split_ms = ms_between(ctx->ev[1], ctx->ev[2]);
    timings->merge_ms = ms_between(ctx->ev[3], ctx->ev[4]);
    timings->total_ms = ms_between(ctx->ev[0], ctx->ev[4]);
  }
  return DRPT_OK;
}

int drpt_autotune(drpt_ctx* ctx, const drpt_params* p, const void* X, int32_t* plan_out) {
  if (!ctx || !X || !plan_out) {
    set_error("autotune: null pointer");
    return DRPT_ERR_ARG;
  }
  DRPT_CHECK(validate(p));
  if (p->precision == DRPT_FP32) {
    set_error("autotune: the FP32 path has a fixed work-item layout (64 rows)");
    return DRPT_ERR_UNSUPPORTED;
  }
  DRPT_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  uint32_t* heap;
  DRPT_CHECK(build_front(ctx, p, X, &heap));
  // the inputs of the pooled filter pass of a build: known lists from the
  // first trees, their K-th distances as every tree's bound
  ItemPlan best = lowp_default_plan();
  LowpRun r;
  DRPT_CHECK(lowp_setup(ctx, p, X, heap, best, &r));
  DRPT_CHECK(lowp_known(ctx, &r, X, best));
  DRPT_CHECK(launch_leaf_gather_val(ctx, r.G.perm, (int64_t)r.T * r.n, r.bnd, 1, 0, r.boundT));
  r.a.known = r.ki;
  const int rb = p->d * 2;  // the pooled passes run in BF16
  const bool ring_ok = rb % 16 == 0 && rb <= 256;
  const bool leaf_ok = pool_leaf_rows(ctx, p->d, kModePool) > 0;  // whole leaves resident in LDS
  hipEvent_t e0 = ctx->ev[10], e1 = ctx->ev[11];
  // candidates per bucket: rows per item x Gram pipeline; each timed twice
  // after a warm-up, the incumbent kept unless another is 3% faster
  const int rpis[4] = {1 << 30, 256, 128, 64};
  for (int b = 0; b < kSizeBuckets; ++b) {
    double best_ms = -1.0;
    for (int ir = 0; ir < 4; ++ir) {
      for (int ring = ir == 0 && leaf_ok ? 2 : 1; ring >= (ring_ok ? 0 : 1); --ring) {
        ItemPlan cand = best;
        cand.rpi[b] = rpis[ir];
        cand.ring[b] = ring;
        DRPT_CHECK(group_impl(ctx, p, X, heap, cand, r.first, &r.G, true));
        double ms = 0.0;
        for (int rep = 0; rep < 3; ++rep) {
          DRPT_CHECK(fill_u32(ctx, (uint32_t*)r.pcnt, 0u, r.n));
          DRPT_CHECK(fill_u32(ctx, (uint32_t*)r.oflag, 0u, r.n));
          DRPT_CHECK(fill_u32(ctx, (uint32_t*)r.novf, 0u, 1));
          DRPT_HIP(hipEventRecord(e0, s));
          DRPT_CHECK(lowp_pass(ctx, &r, kModePool, cand, r.G.range + 4, b));
          DRPT_HIP(hipEventRecord(e1, s));
          DRPT_HIP(hipEventSynchronize(e1));
          if (rep > 0) ms += ms_between(e0, e1);
        }
        if (best_ms < 0.0 || ms < 0.97 * best_ms) {
          best_ms = ms;
          best = cand;
        }
      }
    }
  }
  // trees per pooled filter pass: whole low-precision stages timed per group
  // size (smaller groups: fewer rows leave their pools for the exact path,
  // more merges); unless the caller fixed it
  auto store = [&]() {
    std::vector<int> v(kPlanInts);
    for (int b = 0; b < kSizeBuckets; ++b) {
      v[b] = best.rpi[b];
      v[kSizeBuckets + b] = best.ring[b];
    }
    v[2 * kSizeBuckets] = best.tree_group;
    v[2 * kSizeBuckets + 1] = best.first_trees;
    v[2 * kSizeBuckets + 2] = best.rerank;
    ctx->plans[plan_key(p)] = v;
    for (int i = 0; i < kPlanInts; ++i) plan_out[i] = v[i];
  };
  int32_t* oi;
  float* od;
  DRPT_CHECK(ws(ctx, "tune.out_idx", (size_t)p->n * p->nn, &oi));
  DRPT_CHECK(ws(ctx, "tune.out_dist", (size_t)p->n * p->nn, &od));
  // one low-precision stage with the stored plan (the first run sizes the workspace)
  auto time_stage = [&](double* ms) -> int {
    store();
    for (int rep = 0; rep < 2; ++rep) {
      DRPT_HIP(hipEventRecord(e0, s));
      DRPT_CHECK(graph_knn_lowp(ctx, p, X, heap, oi, od, nullptr));
      DRPT_HIP(hipEventRecord(e1, s));
      DRPT_HIP(hipEventSynchronize(e1));
      if (rep > 0) *ms = ms_between(e0, e1);
    }
    return DRPT_OK;
  };
  const int T = p->ntrees;
  if (p->tree_group == 0) {
    const int cands[7] = {T, 8, 6, 5, 4, 3, 2};
    double best_ms = -1.0;
    int best_g = 0;
    std::vector<int> tried;
    for (int g : cands) {
      if (g < 1 || g > T || std::find(tried.begin(), tried.end(), g) != tried.end()) continue;
      tried.push_back(g);
      best.tree_group = g;
      double ms = 0.0;
      DRPT_CHECK(time_stage(&ms));
      if (best_ms < 0.0 || ms < 0.97 * best_ms) {
        best_ms = ms;
        best_g = g;
      }
    }
    best.tree_group = best_g;
  }
  // trees whose widened candidates give the first bounds: fewer trees, a
  // cheaper candidate pass and first merge, looser bounds for the pools
  {
    const int cands[3] = {kFirstTrees, 2, 1};
    double best_ms = -1.0;
    int best_f = 0;
    for (int f : cands) {
      if (f > T) continue;
      best.first_trees = f;
      double ms = 0.0;
      DRPT_CHECK(time_stage(&ms));
      if (best_ms < 0.0 || ms < 0.97 * best_ms) {
        best_ms = ms;
        best_f = f;
      }
    }
    best.first_trees = best_f;
  }
  // exact re-rank depth of the first bounds: fewer gathers in the first merge,
  // looser bounds for the pools; -1: bounds from approximate distances alone
  {
    const int K = p->nn;
    const int cands[4] = {2 * K, (3 * K + 1) / 2, K, -1};
    double best_ms = -1.0;
    int best_r = 0;
    for (int i = 0; i < 4; ++i) {
      if (i > 0 && cands[i] == cands[i - 1]) continue;
      best.rerank = cands[i];
      double ms = 0.0;
      DRPT_CHECK(time_stage(&ms));
      if (best_ms < 0.0 || ms < 0.97 * best_ms) {
        best_ms = ms;
        best_r = cands[i];
      }
    }
    best.rerank = best_r;
  }
  store();
  DRPT_HIP(hipStreamSynchronize(s));
  return DRPT_OK;
}

int drpt_last_phases(drpt_ctx* ctx, drpt_phases* out) {
  if (!ctx || !out) {
    set_error("last_phases: null pointer");
    return DRPT_ERR_ARG;
  }
  *out = ctx->last_phases;
  return DRPT_OK;
}

int drpt_set_item_plan(drpt_ctx* ctx, const drpt_params* p, const int32_t* plan) {
  if (!ctx || !p || !plan) {
    set_error("set_item_plan: null pointer");
    return DRPT_ERR_ARG;
  }
  DRPT_CHECK(validate(p));
  std::vector<int> v(kPlanInts);
  for (int b = 0; b < kSizeBuckets; ++b) {
    if (plan[b] < 64 || plan[kSizeBuckets + b] < 0 || plan[kSizeBuckets + b] > 2) {
      // the work-item buffer holds ceil(m / 64) items per leaf
      set_error("set_item_plan: rows per item must be >= 64 and ring flags 0, 1 or 2");
      return DRPT_ERR_ARG;
    }
    v[b] = plan[b];
    v[kSizeBuckets + b] = plan[kSizeBuckets + b];
  }
  if (plan[2 * kSizeBuckets] < 0 || plan[2 * kSizeBuckets + 1] < 0 ||
      plan[2 * kSizeBuckets + 2] < -1) {
    set_error("set_item_plan: tree_group and first_trees must be >= 0, rerank >= -1");
    return DRPT_ERR_ARG;
  }
  v[2 * kSizeBuckets] = plan[2 * kSizeBuckets];
  v[2 * kSizeBuckets + 1] = plan[2 * kSizeBuckets + 1];
  v[2 * kSizeBuckets + 2] = plan[2 * kSizeBuckets + 2];
  ctx->plans[plan_key(p)] = v;
  return DRPT_OK;
}

}  // extern "C"
