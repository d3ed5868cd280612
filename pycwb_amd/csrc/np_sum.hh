// numpy's float64 add.reduce order on the device.
//
// np.sum / np.mean of a contiguous float64 array adds 8192-element chunks to
// 0.0 in order; each chunk is summed pairwise (numpy's pairwise_sum,
// loops_utils.h.src): below 8 values a plain loop from 0.0, up to 128 values
// eight strided accumulators combined as ((r0 + r1) + (r2 + r3)) +
// ((r4 + r5) + (r6 + r7)) with the rest added in order, else the sum of the
// two parts split at n / 2 rounded down to a multiple of 8.  The recursion
// runs on an explicit stack; leaves are visited left to right.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pycwb {

constexpr int64_t kNpChunk = 8192;
constexpr int64_t kNpLeaf = 128;

// Sum of g(s), ..., g(s + n - 1) for a leaf (n <= 128).
template <class G>
__device__ double np_leaf(const G& g, int64_t s, int64_t n) {
  if (n < 8) {
    double r = 0.0;
    for (int64_t i = 0; i < n; ++i) r += g(s + i);
    return r;
  }
  double r0 = g(s), r1 = g(s + 1), r2 = g(s + 2), r3 = g(s + 3);
  double r4 = g(s + 4), r5 = g(s + 5), r6 = g(s + 6), r7 = g(s + 7);
  int64_t i = 8;
  for (; i < n - (n % 8); i += 8) {
    r0 += g(s + i);
    r1 += g(s + i + 1);
    r2 += g(s + i + 2);
    r3 += g(s + i + 3);
    r4 += g(s + i + 4);
    r5 += g(s + i + 5);
    r6 += g(s + i + 6);
    r7 += g(s + i + 7);
  }
  double res = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
  for (; i < n; ++i) res += g(s + i);
  return res;
}

// Frames of the explicit recursion stack (depth <= 7 for n <= 8192).
struct NpFrames {
  static constexpr int kDepth = 12;
  int64_t fs[kDepth], fn[kDepth];
  double fv[kDepth];
  int ph[kDepth];
};

// The pairwise tree over [0, n) (n <= 8192): leaf(s, len) gives each leaf's
// value, the tree adds them as numpy does.  F holds the stack: a thread's own
// frames, or frames in LDS for a single thread of a block.
template <class Leaf, class Frames>
__device__ double np_tree_in(int64_t n, Leaf& leaf, Frames& F) {
  int sp = 0;
  F.fs[0] = 0;
  F.fn[0] = n;
  F.ph[0] = 0;
  double ret = 0.0;
  for (;;) {
    if (F.ph[sp] == 0 && F.fn[sp] <= kNpLeaf) {
      ret = leaf(F.fs[sp], F.fn[sp]);
    } else if (F.ph[sp] < 2) {
      int64_t n2 = F.fn[sp] / 2;
      n2 -= n2 % 8;
      if (F.ph[sp] == 1) F.fv[sp] = ret;
      F.fs[sp + 1] = F.ph[sp] == 0 ? F.fs[sp] : F.fs[sp] + n2;
      F.fn[sp + 1] = F.ph[sp] == 0 ? n2 : F.fn[sp] - n2;
      F.ph[sp + 1] = 0;
      F.ph[sp] += 1;
      ++sp;
      continue;
    } else {
      ret = F.fv[sp] + ret;
    }
    if (sp == 0) return ret;
    --sp;
  }
}

template <class Leaf>
__device__ double np_tree(int64_t n, Leaf& leaf) {
  NpFrames F;
  return np_tree_in(n, leaf, F);
}

template <class G>
struct NpLeafSum {
  const G& g;
  int64_t base;
  __device__ double operator()(int64_t s, int64_t n) const { return np_leaf(g, base + s, n); }
};

// np.sum of g(0), ..., g(n - 1), one thread.
template <class G>
__device__ double np_reduce(const G& g, int64_t n) {
  if (n <= kNpLeaf) return 0.0 + np_leaf(g, 0, n);  // one chunk that is one leaf
  double r = 0.0;
  for (int64_t s = 0; s < n; s += kNpChunk) {
    NpLeafSum<G> leaf{g, s};
    r += np_tree(n - s < kNpChunk ? n - s : kNpChunk, leaf);
  }
  return r;
}

}  // namespace pycwb
