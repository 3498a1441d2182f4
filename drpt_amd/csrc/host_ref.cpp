// Host-side pieces whose definition is a libstdc++ routine, compiled with g++
// (as the reference is) so that they produce the reference's values.
//
//  * the projection matrix R: std::mt19937 + uniform_real_distribution<float>
//    + normal_distribution<float> in the order of
//    MathOp::build_sparse_local_random_matrix
//    (cpp/src/drpt/math/math_operations.cpp:48-76);
//  * the reference local tree in its default configuration: MDRPT passes the
//    float density to an int parameter (cpp/src/drpt/algo/mdrpt.hpp:61-62,
//    mdrpt.cpp:117), the local projection is all zero, and DRPTLocal
//    (drpt_local.cpp:121-161) arranges leaf positions with std::nth_element
//    on equal keys, a permutation that depends only on the leaf size;
//  * the local depth rule of mdrpt.cpp:324.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <random>
#include <vector>

namespace drpt {

void host_build_projection(int d, int L, float density, uint32_t seed, float* R) {
  std::mt19937 gen(seed);
  std::uniform_real_distribution<float> uniform(0, 1);
  std::normal_distribution<float> normal(0, 1);
  float* out = R;
  for (int row = 0; row < d; ++row) {
    for (int col = 0; col < L; ++col) {
      // a draw above the density leaves the entry empty; the normal draw is
      // consumed only for kept entries
      float keep = uniform(gen);
      *out++ = keep > density ? 0.0f : static_cast<float>(normal(gen));
    }
  }
}

namespace {

struct EqualKeys {
  bool operator()(int, int) const { return false; }
};

void split_positions(std::vector<int>::iterator first, std::vector<int>::iterator last, int level,
                     int depth) {
  if (level == depth) return;
  const int count = static_cast<int>(last - first);
  std::nth_element(first, first + count / 2, last, EqualKeys());
  auto middle = last - count / 2;
  split_positions(first, middle, level + 1, depth);
  split_positions(middle, last, level + 1, depth);
}

void halving_sizes(int count, int level, int depth, std::vector<int32_t>& out) {
  if (level == depth) {
    out.push_back(count);
    return;
  }
  halving_sizes(count - count / 2, level + 1, depth, out);
  halving_sizes(count / 2, level + 1, depth, out);
}

}  // namespace

void host_local_positions(int m, int ld, int32_t* out) {
  std::vector<int> pos(m);
  std::iota(pos.begin(), pos.end(), 0);
  split_positions(pos.begin(), pos.end(), 0, ld);
  std::copy(pos.begin(), pos.end(), out);
}

void host_local_sizes(int m, int ld, int32_t* out) {
  std::vector<int32_t> sizes;
  halving_sizes(m, 0, ld, sizes);
  std::copy(sizes.begin(), sizes.end(), out);
}

int host_local_depth(int64_t gmin, int nn, int offset) {
  return static_cast<int>(std::log2(static_cast<double>(gmin)) -
                          (std::log2(static_cast<double>(nn)) + offset));
}

}  // namespace drpt
