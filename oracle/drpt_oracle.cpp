// ORACLE — test infrastructure only. Not part of the product.
//
// CPU restatement of HipGraph/DRPT's k-NN graph construction path
// (MDRPT::grow_trees + MDRPT::gather_nns on one rank). Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library:
// as the checker the HIP path is compared against, and as the timed CPU
// baseline. The product (drpt_amd/) never links, loads or calls it.
//
// Each function names the reference code it restates (paths relative to the
// reference repository root). Canonical choices where the reference is not
// reproducible (seeds, OpenMP-order float sums) are marked "CANONICAL".
//
// Pins (what checks each function independently of the HIP path):
//   orc_build_R        mt19937 10000th-output known answer; generate_canonical
//                      and Marsaglia-polar restatement in Python (tests/refpy.py)
//   orc_project        closed form X.R in exact rational arithmetic (Python
//                      fractions) on small inputs; linearity in X
//   orc_node_split     Appendix-A closed form re-derived in Python on small
//                      nodes (brute-force histogram, exact fractions for the
//                      fixed-point sums); strict mode == sequential float32
//   orc_grow_tree      invariants (counts, leaf ranges), brute-force re-run
//                      of orc_node_split per node, monotone-shift invariance
//   orc_local_*        count_leaf_sizes closed form; the known answer m = 8,
//                      ld = 3 -> 6 5 4 7 0 3 2 1; a step-by-step Python
//                      restatement of libstdc++'s introselect (tests/refpy.py)
//                      for every leaf size below 400 and several depths, with
//                      all-equal keys and with real keys (value, index)
//   orc_leaf_knn       brute force over leafmates in Python (FP64 + exact FP32
//                      FMA sequence via numpy float32)
//   orc_merge          set semantics: top-K of the union (Python)
//   orc_bruteforce_f64 exact squared distances on integer data (int64 numpy)
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>
#include <chrono>
#include <omp.h>

extern "C" {

// ----------------------------------------------------------------------------
// R: sparse Gaussian projection matrix, d x L row-major.
// Reference: cpp/src/drpt/math/math_operations.cpp:48-76 (called via :78-85 from
// cpp/src/drpt/algo/mdrpt.cpp:58-59). Column index = tree*D + level.
// libstdc++ <random> is called directly: the reference's R is defined by it.
// ----------------------------------------------------------------------------
void orc_build_R(int d, int L, float density, uint32_t seed, float* R) {
  std::mt19937 gen(seed);
  std::uniform_real_distribution<float> uni(0, 1);
  std::normal_distribution<float> nrm(0, 1);
  for (int j = 0; j < d; ++j)
    for (int i = 0; i < L; ++i)
      R[(size_t)j * L + i] = (uni(gen) > density) ? 0.0f : (float)nrm(gen);
}

// ----------------------------------------------------------------------------
// P = X.R, stored column-major P[c*n + i].
// Reference: math_operations.cpp:15-46 (cblas_sgemm). CANONICAL: each output
// is the double-precision sum of x_ij*r_jc over the non-zeros of column c of R
// in ascending j, rounded once to float (products of u8/f32 by f32 are exact
// in double, so this is order-fixed and FMA-agnostic).
// colmax_bits[c] = max_i bits(|P[c][i]|) (used by the fixed-point statistics).
// ----------------------------------------------------------------------------
static inline double xval(const void* X, int x_is_u8, int d, int64_t i, int j) {
  return x_is_u8 ? (double)((const uint8_t*)X)[(size_t)i * d + j]
                 : (double)((const float*)X)[(size_t)i * d + j];
}

void orc_project(const void* X, int x_is_u8, int64_t n, int d, const float* R, int L,
                 float* P, uint32_t* colmax_bits) {
  std::vector<std::vector<std::pair<int, float>>> nz(L);
  for (int j = 0; j < d; ++j)
    for (int c = 0; c < L; ++c) {
      float w = R[(size_t)j * L + c];
      if (w != 0.0f) nz[c].push_back({j, w});
    }
  for (int c = 0; c < L; ++c) colmax_bits[c] = 0;
#pragma omp parallel
  {
    std::vector<uint32_t> lmax(L, 0);
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      for (int c = 0; c < L; ++c) {
        double acc = 0.0;
        for (auto& e : nz[c]) {
          double prod = xval(X, x_is_u8, d, i, e.first) * (double)e.second;
          acc = acc + prod;
        }
        float p = (float)acc;
        P[(size_t)c * n + i] = p;
        uint32_t b;
        std::memcpy(&b, &p, 4);
        b &= 0x7fffffffu;
        if (b > lmax[c]) lmax[c] = b;
      }
    }
#pragma omp critical
    for (int c = 0; c < L; ++c) colmax_bits[c] = std::max(colmax_bits[c], lmax[c]);
  }
}

// ----------------------------------------------------------------------------
// Fixed-point exponents for the CANONICAL node statistics.
// With e = floor(log2(max|v|)) over the whole column and shift = 63 - bitlen(N):
//   kq1 = e + 1 - shift    (|v| / 2^kq1 <= 2^shift)
//   kq2 = 2(e + 2) - shift (fl(v - mu)^2 / 2^kq2 <= 2^shift)
// so a node sum of <= N terms fits in int64 whatever the order.
// ----------------------------------------------------------------------------
void orc_quant_exponents(uint32_t colmax_bits, int64_t N, int* kq1, int* kq2) {
  int E = (int)((colmax_bits >> 23) & 0xff);
  int e = (E == 0) ? -127 : E - 127;
  int bitlen = 0;
  while (bitlen < 63 && (N >> bitlen) != 0) ++bitlen;
  int shift = 63 - bitlen;
  *kq1 = e + 1 - shift;
  *kq2 = 2 * (e + 2) - shift;
}

static inline int64_t quant(float v, int kq) {
  return (int64_t)std::rint(std::ldexp((double)v, -kq));
}

// ----------------------------------------------------------------------------
// One node's split value: mean, variance, 16-slot edge table, histogram and
// interpolated median. Reference: math_operations.cpp:121-206 (mean/variance)
// and :208-383 (median); the caller's partition rule is drpt_global.cpp:308.
// stat_mode 1 ("strict") = the reference's sequential float32 sums in node
// order (bit-exact with the reference at OMP_NUM_THREADS=1).
// stat_mode 0 (CANONICAL "exact") = order-independent fixed-point sums
// (quantum 2^kq1, 2^kq2), then mu = fl(S/n), var = fl(S2/(n-1)) in double.
// E[-1] (read at :373 when the first slot is selected) is defined as 0, the
// value the previous node's never-written slot 15 holds.
// Returns the median; writes mu, var, E[16], H[16].
// ----------------------------------------------------------------------------
float orc_node_split(const float* v, int64_t n, int stat_mode, int kq1, int kq2,
                     float* mu_out, float* var_out, float* E, uint32_t* H) {
  float mu, var;
  if (stat_mode == 1) {
    float s = 0.0f;
    for (int64_t i = 0; i < n; ++i) s += v[i];
    mu = s / (float)n;
    float q = 0.0f;
    for (int64_t i = 0; i < n; ++i) {
      float diff = v[i] - mu;
      float sq = diff * diff;
      q += sq;
    }
    var = q / (float)(n - 1);
  } else {
    int64_t S = 0;
    for (int64_t i = 0; i < n; ++i) S += quant(v[i], kq1);
    mu = (float)(std::ldexp((double)S, kq1) / (double)n);
    int64_t S2 = 0;
    for (int64_t i = 0; i < n; ++i) {
      float w = v[i] - mu;
      float w2 = w * w;
      S2 += quant(w2, kq2);
    }
    var = (float)(std::ldexp((double)S2, kq2) / (double)(n - 1));
  }
  *mu_out = mu;
  *var_out = var;

  // Edge table: math_operations.cpp:223-288 with no_of_bins = 7
  // (drpt_global.cpp:269): factor 1, 16 slots, steps sigma/30, sigma/6, sigma/2.
  float sigma = std::sqrt(var);
  float s1 = (float)((double)sigma / 30.0);
  float s2 = (float)((double)sigma / 6.0);
  float s3 = (float)((double)sigma / 2.0);
  for (int j = 1; j <= 4; ++j) {
    float r = (float)j * s1;
    E[6 + j] = mu + r;
    E[7 - j] = mu - r;
  }
  for (int j = 1; j <= 2; ++j) {
    float js = (float)j * s2;
    float r = sigma + js;
    E[10 + j] = mu + r;
    E[3 - j] = mu - r;
  }
  {
    float two = 2.0f * sigma;
    float r = two + s3;
    E[13] = mu + r;
    E[0] = mu - r;
  }
  E[14] = 0.0f;
  E[15] = 0.0f;

  // Histogram: math_operations.cpp:292-308 (both tails land in slot 0).
  for (int k = 0; k < 16; ++k) H[k] = 0;
  for (int64_t i = 0; i < n; ++i) {
    float x = v[i];
    int flag = 1;
    for (int j = 1; j < 16; ++j) {
      if (E[j - 1] < x && E[j] >= x) {
        flag = 0;
        H[j] += 1;
      }
    }
    if (flag) H[0] += 1;
  }

  // Cumulative search and interpolation: math_operations.cpp:333-377.
  float cfreq = 0.0f;
  int sel = -1;
  float val = (float)(int64_t)((50 * n) / 100);
  for (int k = 0; k < 16; ++k) {
    cfreq = (float)((double)cfreq + (double)H[k]);
    if (cfreq >= val) {
      sel = k;
      break;
    }
  }
  if (sel < 0) sel = 15;  // unreachable (sum of H == n); defined for safety
  double count = (double)H[sel];
  float lo = (sel == 0) ? 0.0f : E[sel - 1];
  float width = E[sel] - lo;
  double frac = ((double)(n / 2) - ((double)cfreq - count)) / count;
  double med = (double)lo + frac * (double)width;
  return (float)med;
}

// ----------------------------------------------------------------------------
// Grow one tree over P (column-major, column = t*D + level).
// Reference: drpt_global.cpp:153-352 (root = all points in index order,
// heap numbering :231-233/:280-281, "value <= median -> left" :308, children
// keep index order, which is the reference's order at OMP_NUM_THREADS=1).
// mode 0 (fixed, reference): every node at levels 0..D-2 is split, leaves are
//   the 2^(D-1) nodes of depth D-1; leaf id = heap - 2^(D-1) (:283-284).
// mode 1 (leaf-max, this build's policy for large inputs): a node at level
//   l <= D-2 is split iff its size > leaf_max; otherwise it is a leaf.
// node_out[i] = heap index (root = 1) of point i's leaf.
// medians_out (optional, size 2^(D-1)) receives the median of every split
// node by heap index (NaN-free nodes only; unsplit slots untouched).
// ----------------------------------------------------------------------------
int orc_grow_tree(const float* P, int64_t n, int D, int t, int mode, int leaf_max,
                  int stat_mode, const uint32_t* colmax, uint32_t* node_out,
                  float* medians_out) {
  if (D < 1) return -1;
  std::vector<std::vector<int64_t>> cur(1);
  cur[0].resize(n);
  std::iota(cur[0].begin(), cur[0].end(), (int64_t)0);
  std::vector<float> vals;
  for (int l = 0; l < D - 1; ++l) {
    int64_t col = (int64_t)t * D + l;
    int kq1, kq2;
    orc_quant_exponents(colmax[col], n, &kq1, &kq2);
    std::vector<std::vector<int64_t>> nxt(cur.size() * 2);
    for (size_t i = 0; i < cur.size(); ++i) {
      auto& pts = cur[i];
      uint64_t heap = ((uint64_t)1 << l) + i;
      if (pts.empty()) continue;
      bool split = (mode == 0) || ((int64_t)pts.size() > (int64_t)leaf_max);
      if (!split) {
        for (int64_t p : pts) node_out[p] = (uint32_t)heap;
        continue;
      }
      vals.resize(pts.size());
      for (size_t k = 0; k < pts.size(); ++k) vals[k] = P[(size_t)col * n + pts[k]];
      float mu, var, E[16];
      uint32_t H[16];
      float med = orc_node_split(vals.data(), (int64_t)pts.size(), stat_mode, kq1, kq2,
                                 &mu, &var, E, H);
      if (medians_out) medians_out[heap] = med;
      auto& L = nxt[2 * i];
      auto& Rr = nxt[2 * i + 1];
      for (size_t k = 0; k < pts.size(); ++k) {
        if (vals[k] <= med) L.push_back(pts[k]);
        else Rr.push_back(pts[k]);
      }
    }
    cur.swap(nxt);
  }
  uint64_t base = (uint64_t)1 << (D - 1);
  for (size_t i = 0; i < cur.size(); ++i)
    for (int64_t p : cur[i]) node_out[p] = (uint32_t)(base + i);
  return 0;
}

// ----------------------------------------------------------------------------
// Local stage in the reference's default configuration.
// Reference: mdrpt.cpp:321-400 passes the float density to an int parameter
// (mdrpt.hpp:61-62), so the local projection is the R-generator with density
// 0 (math_operations.cpp:64: kept only if a uniform draw is exactly 0), and
// the local projected values are all zero. drpt_local.cpp:121-161 then splits
// positions with std::nth_element on equal keys; the leaves are positions
// [first[i], first[i+1]) of the permuted index array (:163-180) with sizes
// from the ceil/floor halving of :49-69.
// orc_local_perm writes the permuted positions (size m);
// orc_local_sizes writes the 2^ld leaf sizes.
// ----------------------------------------------------------------------------
static void local_recurse(std::vector<int>::iterator b, std::vector<int>::iterator e,
                          int depth, int ld) {
  int size = (int)(e - b);
  if (depth == ld) return;
  std::nth_element(b, b + size / 2, e, [](int, int) { return 0.0f < 0.0f; });
  auto mid = e - size / 2;
  local_recurse(b, mid, depth + 1, ld);
  local_recurse(mid, e, depth + 1, ld);
}

void orc_local_perm(int m, int ld, int32_t* out) {
  std::vector<int> idx(m);
  std::iota(idx.begin(), idx.end(), 0);
  local_recurse(idx.begin(), idx.end(), 0, ld);
  for (int i = 0; i < m; ++i) out[i] = idx[i];
}

static void local_sizes_rec(int size, int level, int depth, std::vector<int>& out) {
  if (level == depth) {
    out.push_back(size);
    return;
  }
  local_sizes_rec(size - size / 2, level + 1, depth, out);
  local_sizes_rec(size / 2, level + 1, depth, out);
}

void orc_local_sizes(int m, int ld, int32_t* out) {
  std::vector<int> s;
  local_sizes_rec(m, 0, ld, s);
  for (size_t i = 0; i < s.size(); ++i) out[i] = s[i];
}

// ----------------------------------------------------------------------------
// True local random projection: with -density k >= 1 (main.cpp:69-72 parses
// an integer) the int density handed to grow_local_trees is k >= 1
// (mdrpt.cpp:321-400), so the local matrix C (d x ld, math_operations.cpp:48-76
// with seed S+1+t) is non-zero and the local tree splits on real keys.
// Reference: drpt_local.cpp:121-161 -- at depth l a node of s positions is
// rearranged by std::nth_element(begin, begin + s/2, end, less-by-LP[l]) and
// its left child is [begin, end - s/2): the ceil(s/2) smallest values; the
// right child the rest; the left child is recursed first.
// CANONICAL: ties in LP are broken by point index (the reference leaves them
// to libstdc++'s introselect), and each node is fully sorted (only the sets
// matter downstream). LP is column-major [ld][n] (P of the local matrix),
// pts the m points of one global leaf; on return each local leaf is
// contiguous, in the order and sizes of orc_local_sizes.
// ----------------------------------------------------------------------------
static void local_split_rec(const float* LP, int64_t n, int32_t* b, int32_t* e, int depth,
                            int ld) {
  if (depth == ld) return;
  const float* v = LP + (size_t)depth * n;
  std::sort(b, e, [v](int32_t x, int32_t y) {
    if (v[x] < v[y]) return true;
    if (v[y] < v[x]) return false;
    return x < y;
  });
  const int64_t size = e - b;
  int32_t* mid = e - size / 2;
  local_split_rec(LP, n, b, mid, depth + 1, ld);
  local_split_rec(LP, n, mid, e, depth + 1, ld);
}

void orc_local_split(const float* LP, int64_t n, int ld, int32_t* pts, int m) {
  local_split_rec(LP, n, pts, pts + m, 0, ld);
}

// Local matrix check: the reference generates C with the truncated density
// (mdrpt.cpp:348-352); returns the number of non-zeros (0 in practice).
int orc_local_C_nonzeros(int d, int ld, float density_int, uint32_t seed) {
  std::vector<float> C((size_t)d * ld);
  orc_build_R(d, ld, density_int, seed, C.data());
  int nz = 0;
  for (float c : C) nz += (c != 0.0f);
  return nz;
}

// ----------------------------------------------------------------------------
// CANONICAL exact distance: D(x,y) = fl-sequence acc = fma(t_k, t_k, acc),
// t_k = fl(x_k - y_k), k ascending. Reference: math_operations.cpp:385-396
// (saxpy + snrm2, i.e. the unsquared norm with BLAS-specific rounding); the
// squared, order-fixed form ranks identically up to ties.
// ----------------------------------------------------------------------------
static inline float xf(const void* X, int x_is_u8, int d, int64_t i, int k) {
  return x_is_u8 ? (float)((const uint8_t*)X)[(size_t)i * d + k]
                 : ((const float*)X)[(size_t)i * d + k];
}

float orc_dist(const void* X, int x_is_u8, int d, int64_t i, int64_t j) {
  float acc = 0.0f;
  for (int k = 0; k < d; ++k) {
    float t = xf(X, x_is_u8, d, i, k) - xf(X, x_is_u8, d, j, k);
    acc = std::fma(t, t, acc);
  }
  return acc;
}

// Leaf-block distances, vectorised across 8 columns; the per-pair summation
// order is the canonical one (k ascending, one fma per k). The AVX2+FMA body
// and the portable body compute the same values (fma is exact either way).
#define DIST_ROWS_BODY                                              \
  int j = 0;                                                        \
  for (; j + 8 <= m; j += 8) {                                      \
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};                        \
    for (int k = 0; k < d; ++k) {                                   \
      float a = xi[k];                                              \
      const float* col = xT + (size_t)k * m + j;                    \
      for (int u = 0; u < 8; ++u) {                                 \
        float t = a - col[u];                                       \
        acc[u] = std::fma(t, t, acc[u]);                            \
      }                                                             \
    }                                                               \
    for (int u = 0; u < 8; ++u) out[j + u] = acc[u];                \
  }                                                                 \
  for (; j < m; ++j) {                                              \
    float acc = 0.0f;                                               \
    for (int k = 0; k < d; ++k) {                                   \
      float t = xi[k] - xT[(size_t)k * m + j];                      \
      acc = std::fma(t, t, acc);                                    \
    }                                                               \
    out[j] = acc;                                                   \
  }

__attribute__((target("avx2,fma")))
static void dist_rows_fma(const float* xi, const float* xT, int m, int d, float* out) {
  DIST_ROWS_BODY
}
static void dist_rows_plain(const float* xi, const float* xT, int m, int d, float* out) {
  DIST_ROWS_BODY
}
static void dist_rows(const float* xi, const float* xT, int m, int d, float* out) {
  static const bool has_fma = __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");
  if (has_fma) dist_rows_fma(xi, xT, m, d, out);
  else dist_rows_plain(xi, xT, m, d, out);
}

// ----------------------------------------------------------------------------
// Leaf k-NN for one tree. Reference: mdrpt.cpp:123-212 — every ordered pair
// of leafmates, per-row sort, keep the head. CANONICAL: rows are ordered by
// (D, index), self is excluded, K = nn non-self entries are kept (the
// reference keeps 2*nn including self; per-tree truncation at >= nn non-self
// entries is lossless for the final nn-list). Short rows are padded with
// (-1, +inf).
// perm: point ids grouped by leaf; leaf_off[nleaves+1].
// ----------------------------------------------------------------------------
void orc_leaf_knn(const void* X, int x_is_u8, int64_t n, int d, const int32_t* perm,
                  const int64_t* leaf_off, int64_t nleaves, int K, int32_t* cand_idx,
                  float* cand_dist) {
#pragma omp parallel
  {
    std::vector<float> xT, xi(d), row;
    std::vector<int> order;
#pragma omp for schedule(dynamic, 1)
    for (int64_t lf = 0; lf < nleaves; ++lf) {
      int64_t o = leaf_off[lf];
      int m = (int)(leaf_off[lf + 1] - o);
      if (m <= 0) continue;
      xT.assign((size_t)d * m, 0.0f);
      for (int j = 0; j < m; ++j)
        for (int k = 0; k < d; ++k) xT[(size_t)k * m + j] = xf(X, x_is_u8, d, perm[o + j], k);
      row.resize(m);
      order.resize(m);
      for (int a = 0; a < m; ++a) {
        int64_t pa = perm[o + a];
        for (int k = 0; k < d; ++k) xi[k] = xT[(size_t)k * m + a];
        dist_rows(xi.data(), xT.data(), m, d, row.data());
        std::iota(order.begin(), order.end(), 0);
        std::sort(order.begin(), order.end(), [&](int u, int v) {
          if (row[u] != row[v]) return row[u] < row[v];
          return perm[o + u] < perm[o + v];
        });
        int w = 0;
        for (int s = 0; s < m && w < K; ++s) {
          int u = order[s];
          if (perm[o + u] == pa) continue;
          cand_idx[(size_t)pa * K + w] = perm[o + u];
          cand_dist[(size_t)pa * K + w] = row[u];
          ++w;
        }
        for (; w < K; ++w) {
          cand_idx[(size_t)pa * K + w] = -1;
          cand_dist[(size_t)pa * K + w] = INFINITY;
        }
      }
    }
  }
}

// ----------------------------------------------------------------------------
// Cross-tree merge. Reference: mdrpt.cpp:180-198 (std::merge by distance,
// adjacent unique by index, no truncation) and the writer's self filter
// (file_writer.cpp:54-58). CANONICAL: order (D, index) and a full dedup (the
// same index always carries the same D, so equal indices are adjacent), first
// K kept. cand_* are [T][n][K].
// ----------------------------------------------------------------------------
void orc_merge(int64_t n, int T, int K, const int32_t* cand_idx, const float* cand_dist,
               int32_t* out_idx, float* out_dist) {
#pragma omp parallel
  {
    std::vector<std::pair<float, int32_t>> all;
#pragma omp for schedule(static)
    for (int64_t p = 0; p < n; ++p) {
      all.clear();
      for (int t = 0; t < T; ++t)
        for (int k = 0; k < K; ++k) {
          size_t o = ((size_t)t * n + p) * K + k;
          if (cand_idx[o] >= 0) all.push_back({cand_dist[o], cand_idx[o]});
        }
      std::sort(all.begin(), all.end());
      int w = 0;
      int32_t last = -1;
      for (auto& e : all) {
        if (w >= K) break;
        if (e.second == last) continue;
        last = e.second;
        out_idx[(size_t)p * K + w] = e.second;
        out_dist[(size_t)p * K + w] = e.first;
        ++w;
      }
      for (; w < K; ++w) {
        out_idx[(size_t)p * K + w] = -1;
        out_dist[(size_t)p * K + w] = INFINITY;
      }
    }
  }
}

// ----------------------------------------------------------------------------
// FP64 brute-force k-NN for recall: non-self neighbours ordered by
// (FP64 squared distance, index).
// ----------------------------------------------------------------------------
void orc_bruteforce_f64(const void* X, int x_is_u8, int64_t n, int d, const int64_t* queries,
                        int64_t nq, int K, int32_t* out_idx, double* out_d) {
#pragma omp parallel
  {
    std::vector<std::pair<double, int64_t>> best;
#pragma omp for schedule(dynamic, 4)
    for (int64_t qi = 0; qi < nq; ++qi) {
      int64_t q = queries[qi];
      best.clear();
      for (int64_t j = 0; j < n; ++j) {
        if (j == q) continue;
        double s = 0.0;
        for (int k = 0; k < d; ++k) {
          double t = xval(X, x_is_u8, d, q, k) - xval(X, x_is_u8, d, j, k);
          s += t * t;
        }
        if ((int)best.size() < K) {
          best.push_back({s, j});
          std::push_heap(best.begin(), best.end());
        } else if (std::make_pair(s, j) < best.front()) {
          std::pop_heap(best.begin(), best.end());
          best.back() = {s, j};
          std::push_heap(best.begin(), best.end());
        }
      }
      std::sort(best.begin(), best.end());
      for (int k = 0; k < K; ++k) {
        out_idx[qi * K + k] = k < (int)best.size() ? (int32_t)best[k].second : -1;
        out_d[qi * K + k] = k < (int)best.size() ? best[k].first : INFINITY;
      }
    }
  }
}

// ----------------------------------------------------------------------------
// Parameter derivation. Reference: main.cpp:129-132 (density), :141-145
// (tree_depth), mdrpt.cpp:52 (global depth), mdrpt.cpp:324 (local depth).
// ----------------------------------------------------------------------------
int orc_tree_depth(int64_t n) { return (int)(std::log2((double)n) - 2); }
int orc_global_depth(int tree_depth, double ratio) { return (int)(tree_depth * ratio); }
int orc_local_depth(int64_t gmin, int nn, int offset) {
  return (int)(std::log2((double)gmin) - (std::log2((double)nn) + offset));
}

// ----------------------------------------------------------------------------
// Whole build (the CPU baseline and end-to-end checker).
// ----------------------------------------------------------------------------
typedef struct {
  int64_t n;
  int d;
  int x_is_u8;
  int ntrees;
  int depth;        // columns per tree: Dg (fixed mode) or max depth (leaf-max mode)
  int nn;           // K non-self neighbours kept per point
  int local_mode;   // 0 = reference local stage (fixed global tree), 1 = leaf-max
  int leaf_max;
  int local_offset; // reference -local-tree-offset
  int stat_mode;    // 0 canonical exact, 1 strict sequential float
  uint32_t seed;
  float density;
  int nthreads;
  int pad;
} orc_params;

// timings[0..4] = projection, trees, leaves/local, leaf k-NN, merge (seconds).
int orc_build_graph(const orc_params* pp, const void* X, int32_t* out_idx, float* out_dist,
                    double* timings, uint32_t* leaf_heap_out /* optional [T][n] */) {
  const orc_params& p = *pp;
  if (p.nthreads > 0) omp_set_num_threads(p.nthreads);
  auto now = [] { return std::chrono::steady_clock::now(); };
  auto secs = [](auto a, auto b) { return std::chrono::duration<double>(b - a).count(); };
  int T = p.ntrees, D = p.depth, K = p.nn;
  int64_t n = p.n;
  int L = T * D;
  auto t0 = now();
  std::vector<float> R((size_t)p.d * L);
  orc_build_R(p.d, L, p.density, p.seed, R.data());
  std::vector<float> P((size_t)L * n);
  std::vector<uint32_t> colmax(L);
  orc_project(X, p.x_is_u8, n, p.d, R.data(), L, P.data(), colmax.data());
  auto t1 = now();
  std::vector<uint32_t> heap((size_t)T * n);
#pragma omp parallel for schedule(dynamic, 1)
  for (int t = 0; t < T; ++t)
    orc_grow_tree(P.data(), n, D, t, p.local_mode == 0 ? 0 : 1, p.leaf_max, p.stat_mode,
                  colmax.data(), heap.data() + (size_t)t * n, nullptr);
  if (leaf_heap_out) std::memcpy(leaf_heap_out, heap.data(), sizeof(uint32_t) * T * n);
  auto t2 = now();
  // group by leaf (ascending heap index, ascending point id inside)
  std::vector<std::vector<int32_t>> perms(T);
  std::vector<std::vector<int64_t>> offs(T);
  for (int t = 0; t < T; ++t) {
    const uint32_t* h = heap.data() + (size_t)t * n;
    std::vector<int32_t>& pm = perms[t];
    pm.resize(n);
    std::iota(pm.begin(), pm.end(), 0);
    std::stable_sort(pm.begin(), pm.end(), [&](int32_t a, int32_t b) { return h[a] < h[b]; });
    std::vector<int64_t>& of = offs[t];
    of.clear();
    for (int64_t i = 0; i < n; ++i)
      if (i == 0 || h[pm[i]] != h[pm[i - 1]]) of.push_back(i);
    of.push_back(n);
  }
  if (p.local_mode == 0) {
    int64_t gmin = INT64_MAX;
    for (int t = 0; t < T; ++t)
      for (size_t i = 0; i + 1 < offs[t].size(); ++i)
        gmin = std::min(gmin, offs[t][i + 1] - offs[t][i]);
    int ld = orc_local_depth(gmin, K, p.local_offset);
    if (ld <= 0) return -2;  // reference throws (drpt_local.cpp:84-87)
    float dint = (float)(int)p.density;
    int nl = 1 << ld;
    std::vector<float> C((size_t)p.d * ld), LP, Rt;
    std::vector<uint32_t> lcm(ld);
    for (int t = 0; t < T; ++t) {
      // the tree's local matrix: all zero (density < 1 truncated to 0) gives
      // the libstdc++ tie permutation, non-zero the real-key local tree
      orc_build_R(p.d, ld, dint, p.seed + 1u + (uint32_t)t, C.data());
      bool real = false;
      for (float c : C) real = real || c != 0.0f;
      if (real) {
        LP.resize((size_t)ld * n);
        orc_project(X, p.x_is_u8, n, p.d, C.data(), ld, LP.data(), lcm.data());
      }
      std::vector<int32_t> npm(n);
      std::vector<int64_t> nof;
      std::vector<int32_t> tbl, sz(nl);
      for (size_t g = 0; g + 1 < offs[t].size(); ++g) {
        int64_t o = offs[t][g];
        int m = (int)(offs[t][g + 1] - o);
        orc_local_sizes(m, ld, sz.data());
        if (real) {
          for (int k = 0; k < m; ++k) npm[o + k] = perms[t][o + k];
          orc_local_split(LP.data(), n, ld, npm.data() + o, m);
        } else {
          tbl.resize(m);
          orc_local_perm(m, ld, tbl.data());
          for (int k = 0; k < m; ++k) npm[o + k] = perms[t][o + tbl[k]];
        }
        int64_t acc = o;
        for (int s = 0; s < nl; ++s) {
          if (sz[s] > 0) nof.push_back(acc);
          acc += sz[s];
        }
      }
      nof.push_back(n);
      perms[t].swap(npm);
      offs[t].swap(nof);
    }
  }
  auto t3 = now();
  std::vector<int32_t> cidx((size_t)T * n * K);
  std::vector<float> cdist((size_t)T * n * K);
  for (int t = 0; t < T; ++t)
    orc_leaf_knn(X, p.x_is_u8, n, p.d, perms[t].data(), offs[t].data(),
                 (int64_t)offs[t].size() - 1, K, cidx.data() + (size_t)t * n * K,
                 cdist.data() + (size_t)t * n * K);
  auto t4 = now();
  orc_merge(n, T, K, cidx.data(), cdist.data(), out_idx, out_dist);
  auto t5 = now();
  if (timings) {
    timings[0] = secs(t0, t1);
    timings[1] = secs(t1, t2);
    timings[2] = secs(t2, t3);
    timings[3] = secs(t3, t4);
    timings[4] = secs(t4, t5);
  }
  return 0;
}

}  // extern "C"
