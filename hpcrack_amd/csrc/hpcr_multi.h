// Multi-problem launches: one grid runs the same kernel body for up to
// MULTI_MAX independent problems (contexts), each over its own block range
// with its own argument record.  A kernel type K provides
//   Args                 its argument record (trivially copyable)
//   NT, MINW             threads per block and the launch-bounds occupancy
//   block()              the block shape
//   run(args, blk)       the body, given the problem's own block index
// Problem ranges start at multiples of 8 blocks, so that a problem's blocks
// map to the 8 XCDs as in a launch of its own; the padding blocks exit.
#pragma once
#include <algorithm>
#include <cstring>
#include <type_traits>

#include "hpcr_internal.h"
#include "hpcr_tile.h"

namespace hpcr {

constexpr int MULTI_MAX = 8;

template <class K>
struct MultiArgs {
  int n;
  unsigned start[MULTI_MAX + 1];
  unsigned gx[MULTI_MAX], gy[MULTI_MAX], gz[MULTI_MAX];
  typename K::Args a[MULTI_MAX];
};

template <class K>
__global__ __launch_bounds__(K::NT, K::MINW) void multi_kernel(const MultiArgs<K> m) {
  const unsigned b = blockIdx.x;
  int p = 0;
  while (p + 1 < m.n && b >= m.start[p + 1]) p++;
  const unsigned l = b - m.start[p];
  const unsigned gx = m.gx[p], gy = m.gy[p], gxy = gx * gy;
  if (l >= gxy * m.gz[p]) return;  // padding block
  const Blk bk{l % gx, (l / gx) % gy, l / gxy, gx, gy, m.gz[p]};
  // the problem's record as a value, as a kernel of its own receives it
  const typename K::Args a = m.a[p];
  K::run(a, bk);
}

template <class K>
void multi_launch(hipStream_t s, const Rec* const* r, int n) {
  using A = typename K::Args;
  static_assert(sizeof(MultiArgs<K>) <= 4096, "multi-problem kernel arguments exceed 4 KiB");
  for (int i0 = 0; i0 < n; i0 += MULTI_MAX) {
    MultiArgs<K> m;
    std::memset(static_cast<void*>(&m), 0, sizeof m);
    m.n = std::min(MULTI_MAX, n - i0);
    unsigned b = 0;
    for (int p = 0; p < m.n; p++) {
      const Rec& q = *r[i0 + p];
      m.start[p] = b;
      m.gx[p] = q.g[0];
      m.gy[p] = q.g[1];
      m.gz[p] = q.g[2];
      std::memcpy(static_cast<void*>(&m.a[p]), q.a, sizeof(A));
      b += (q.g[0] * q.g[1] * q.g[2] + 7u) & ~7u;
    }
    for (int p = m.n; p <= MULTI_MAX; p++) m.start[p] = b;
    hipLaunchKernelGGL(multi_kernel<K>, dim3(b), K::block(), 0, s, m);
  }
}

// Record the launch (grid g, arguments a) when the calling thread has a
// recorder; false: the caller launches it.
template <class K>
bool record(dim3 g, const typename K::Args& a) {
  Recorder* R = t_recorder;
  if (!R) return false;
  using A = typename K::Args;
  static_assert(sizeof(A) <= REC_ARG_BYTES, "argument record too large");
  static_assert(std::is_trivially_copyable<A>::value, "argument records are copied bytewise");
  Rec r;
  r.fn = &multi_launch<K>;
  r.g[0] = g.x;
  r.g[1] = g.y;
  r.g[2] = g.z;
  std::memset(r.a, 0, sizeof r.a);
  std::memcpy(r.a, static_cast<const void*>(&a), sizeof(A));
  R->recs.push_back(r);
  return true;
}

// tile kernels (hpcr_tile.h) as multi-problem kernel types
template <class Op>
struct TileK {
  struct Args {
    LevelGeom L;
    Op op;
  };
  static constexpr int NT = TPB, MINW = Op::MIN_WAVES;
  static dim3 block() { return dim3(TPB); }
  __device__ __forceinline__ static void run(const Args& a, const Blk& bk) { tile_body<Op>(a.L, a.op, bk); }
};
template <class Op>
struct TileSplitK {
  struct Args {
    LevelGeom L;
    Op op;
    SplitW S;
  };
  static constexpr int NT = TPB, MINW = Op::MIN_WAVES;
  static dim3 block() { return dim3(TPB); }
  __device__ __forceinline__ static void run(const Args& a, const Blk& bk) {
    tile_body_split<Op>(a.L, a.op, a.S, bk);
  }
};

}  // namespace hpcr
