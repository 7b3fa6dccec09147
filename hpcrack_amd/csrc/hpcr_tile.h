// Device building blocks shared by the FP64 kernels (hpcr_kernels.hip) and
// the FP32 multigrid preconditioner (hpcr_mixed.hip): index helpers,
// deterministic reductions, the sum-factorised hex8 element algebra
// (templated on the arithmetic type), the Dirichlet mask and the z-marching
// tile kernel framework.  Included by .hip sources only.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "hpcr_internal.h"

namespace hpcr {

// L_i(g): 1-D linear shape function of node i at Gauss point g (xi = -+1/sqrt3)
#define GA 0.78867513459481288225  // (1 + 1/sqrt(3)) / 2
#define GB 0.21132486540518711775  // (1 - 1/sqrt(3)) / 2

__device__ __forceinline__ double Lw(int i, int g) { return (i == g) ? GA : GB; }

__device__ __forceinline__ size_t nidx(const LevelGeom& L, int i, int j, int k) {
  return (size_t)i + (size_t)L.nx * ((size_t)j + (size_t)L.ny * (size_t)k);
}
__host__ __device__ __forceinline__ size_t Nof(const LevelGeom& L) { return (size_t)L.nx * L.ny * L.nz; }

// Full-weighting restriction stencil at one coarse node (i, j, k)
// (coarsen_f_standard_routine, multigrid_intergrid.cpp:124-342) in its
// separable form: 1-D weights 1/2, 1, 1/2 per axis around fine node
// (2i, 2j, 2k), out-of-range fine neighbours weighted 0.  The 27 loads of a
// component are unconditional, from indices clamped to the grid, so they are
// all in flight together; an out-of-range term then adds 0 * r, which leaves
// each partial sum as it is (up to the sign of an exact zero), so the value is
// that of the form that skips those terms.
template <class P>
struct Restrict27 {
  size_t row[3][3];  // fine node index of (0, y, z) for each (dz, dy)
  int xs[3];
  P wx[3], wy[3], wz[3];
  __device__ __forceinline__ Restrict27(const LevelGeom& Lf, int i, int j, int k) {
    const int x = 2 * i, y = 2 * j, z = 2 * k;
    int ys[3], zs[3];
#pragma unroll
    for (int t = 0; t < 3; t++) {
      const bool ox = x + t - 1 < 0 || x + t - 1 > Lf.nx - 1;
      const bool oy = y + t - 1 < 0 || y + t - 1 > Lf.ny - 1;
      const bool oz = z + t - 1 < 0 || z + t - 1 > Lf.nz - 1;
      wx[t] = ox ? P(0) : (t == 1 ? P(1) : P(0.5));
      wy[t] = oy ? P(0) : (t == 1 ? P(1) : P(0.5));
      wz[t] = oz ? P(0) : (t == 1 ? P(1) : P(0.5));
      xs[t] = min(max(x + t - 1, 0), Lf.nx - 1);
      ys[t] = min(max(y + t - 1, 0), Lf.ny - 1);
      zs[t] = min(max(z + t - 1, 0), Lf.nz - 1);
    }
#pragma unroll
    for (int dz = 0; dz < 3; dz++)
#pragma unroll
      for (int dy = 0; dy < 3; dy++) row[dz][dy] = nidx(Lf, 0, ys[dy], zs[dz]);
  }
  // sum_z wz (sum_y wy (sum_x wx r)) of one component
  __device__ __forceinline__ P operator()(const P* __restrict__ r) const {
    P v[3][3][3];
#pragma unroll
    for (int dz = 0; dz < 3; dz++)
#pragma unroll
      for (int dy = 0; dy < 3; dy++)
#pragma unroll
        for (int dx = 0; dx < 3; dx++) v[dz][dy][dx] = r[row[dz][dy] + xs[dx]];
    P acc = P(0);
#pragma unroll
    for (int dz = 0; dz < 3; dz++) {
      P accy = P(0);
#pragma unroll
      for (int dy = 0; dy < 3; dy++) {
        P accx = P(0);
#pragma unroll
        for (int dx = 0; dx < 3; dx++) accx += wx[dx] * v[dz][dy][dx];
        accy += wy[dy] * accx;
      }
      acc += wz[dz] * accy;
    }
    return acc;
  }
};

// Node index as a 32-bit count, and loads/stores at a 32-bit byte offset from
// a kernel-uniform base pointer: the instruction takes the base from scalar
// registers and the offset from one vector register, so an access costs no
// 64-bit address arithmetic.  A level holds at most MAX_LEVEL_NODES nodes
// (checked when a context is created), so the byte offset of a node within
// one component of a field fits in 32 bits.
__device__ __forceinline__ unsigned nid32(const LevelGeom& L, int i, int j, int k) {
  return (unsigned)i + (unsigned)L.nx * ((unsigned)j + (unsigned)L.ny * (unsigned)k);
}
template <class T>
__device__ __forceinline__ T ldo(const T* base, unsigned boff) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + boff);
}
template <class T>
__device__ __forceinline__ void sto(T* base, unsigned boff, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + boff) = v;
}

// Trilinear prolongation of a coarse correction (multigrid_intergrid.cpp:416-520)
// in its separable form, for the two fine nodes (2i, y, z) and (2i+1, y, z) of
// one thread: the coarse values at the corners of each node's coarse cell,
// weight 1/2 on each odd axis.  The odd node's corners are coarse columns i
// and i+1, the even node's column i twice, so 8 loads per component serve
// both nodes.  Every term enters each node's sum: one that the sum does not
// take (the second column of an even node, the upper corner on an even y or
// z) is multiplied by 0 and added, which leaves the sum as it is (up to the
// sign of an exact zero), so the sum's order and rounding are those of the
// form that skips the terms.  has1 false: the row ends at the even node,
// whose column-(i+1) loads then read column i again.
template <class P>
struct ProlongPair {
  unsigned o[2][2][2];  // coarse node index of [dz][dy][column i, i+1]
  P fy, fz, w0, w1;     // weights of the even and the odd node
  __device__ __forceinline__ ProlongPair(const LevelGeom& Lc, int i, bool has1, int y, int z) {
    const int j0 = y >> 1, k0 = z >> 1;
    const int i1 = has1 ? i + 1 : i, j1 = j0 + (y & 1), k1 = k0 + (z & 1);
#pragma unroll
    for (int dz = 0; dz < 2; dz++)
#pragma unroll
      for (int dy = 0; dy < 2; dy++)
#pragma unroll
        for (int dx = 0; dx < 2; dx++) o[dz][dy][dx] = nid32(Lc, dx ? i1 : i, dy ? j1 : j0, dz ? k1 : k0);
    fy = (y & 1) ? P(1) : P(0);
    fz = (z & 1) ? P(1) : P(0);
    const P wy = (y & 1) ? P(0.5) : P(1), wz = (z & 1) ? P(0.5) : P(1);
    w0 = (P(1) * wy) * wz;
    w1 = (P(0.5) * wy) * wz;
  }
  // the corrections of the even (c0) and the odd (c1) node from one
  // component of the coarse field e
  __device__ __forceinline__ void operator()(const P* __restrict__ e, P& c0, P& c1) const {
    P v[2][2][2];
#pragma unroll
    for (int dz = 0; dz < 2; dz++)
#pragma unroll
      for (int dy = 0; dy < 2; dy++)
#pragma unroll
        for (int dx = 0; dx < 2; dx++) v[dz][dy][dx] = e[o[dz][dy][dx]];
    auto node = [&](P fx, int cx) {
      const P a00 = v[0][0][0] + fx * v[0][0][cx];
      const P a10 = fy * (v[0][1][0] + fx * v[0][1][cx]);
      const P a01 = fz * (v[1][0][0] + fx * v[1][0][cx]);
      const P a11 = (fy * fz) * (v[1][1][0] + fx * v[1][1][cx]);
      return a00 + a10 + a01 + a11;
    };
    c0 = w0 * node(P(0), 0);
    c1 = w1 * node(P(1), 1);
  }
};
// threads of a pairwise prolongation onto level Lf, and the pair of thread q:
// even node x = 2 ip of row (y, z), the odd one x + 1 when has1
__host__ __device__ __forceinline__ size_t prolong_pairs(const LevelGeom& Lf) {
  return (size_t)((Lf.nx + 1) / 2) * Lf.ny * Lf.nz;
}
__device__ __forceinline__ void pair_ijk(const LevelGeom& Lf, unsigned q, int& ip, bool& has1, int& y, int& z) {
  const unsigned px = (unsigned)(Lf.nx + 1) / 2;
  const unsigned row = q / px;
  ip = (int)(q - row * px);
  z = (int)(row / (unsigned)Lf.ny);
  y = (int)(row - (unsigned)z * (unsigned)Lf.ny);
  has1 = 2 * ip + 1 < Lf.nx;
}

// Loads and stores through buffer resources.  The instruction takes the
// base from the resource in scalar registers and a 32-bit byte offset from one
// vector register, so no per-lane 64-bit address is formed, and the hardware's
// range check makes a load at or past the resource's size return 0: a fetch of
// a node outside the grid (offset BOFF_NONE) needs no branch and no zeroing.
// A resource covers one component of one field of one level, at most
// MAX_LEVEL_NODES * 8 <= BOFF_NONE bytes.
constexpr unsigned BOFF_NONE = 0xFFFFFFF8u;
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t brsrc(const void* base, size_t bytes) {
  // dword 3: 32-bit data format, the raw-buffer setting of gfx9 parts
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)(unsigned)bytes, 0x00020000);
}
__device__ __forceinline__ double bld(__amdgpu_buffer_rsrc_t r, unsigned boff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, boff, 0, 0));
}
__device__ __forceinline__ float bldf(__amdgpu_buffer_rsrc_t r, unsigned boff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, boff, 0, 0));
}
template <class P>
__device__ __forceinline__ P bldp(__amdgpu_buffer_rsrc_t r, unsigned boff) {
  if constexpr (std::is_same<P, double>::value)
    return bld(r, boff);
  else
    return bldf(r, boff);
}
__device__ __forceinline__ signed char bld_i8(__amdgpu_buffer_rsrc_t r, unsigned boff) {
  return (signed char)__builtin_amdgcn_raw_buffer_load_b8(r, boff, 0, 0);
}
__device__ __forceinline__ void bst(__amdgpu_buffer_rsrc_t r, unsigned boff, double v) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2_t, v), r, boff, 0, 0);
}

// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------
enum RedOp { RED_SUM = 0, RED_MIN = 1, RED_MAX = 2 };

template <int OP>
__device__ __forceinline__ double red2(double a, double b) {
  if (OP == RED_SUM) return a + b;
  if (OP == RED_MIN) return (b < a) ? b : a;
  return (b > a) ? b : a;
}

template <int OP>
__device__ __forceinline__ double wave_red(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = red2<OP>(v, __shfl_down(v, off, 64));
  return v;
}

// block of blockDim.x (multiple of 64, <= 1024) threads -> result in thread 0
template <int OP, int NV>
__device__ __forceinline__ void block_red(double (&v)[NV], double* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = wave_red<OP>(v[k]);
  if (lane == 0)
    for (int k = 0; k < NV; k++) sh[k * 16 + wid] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 0; k < NV; k++) {
      double a = sh[k * 16];
      for (int w = 1; w < nw; w++) a = red2<OP>(a, sh[k * 16 + w]);
      v[k] = a;
    }
}

// Block index and grid of a kernel body: the hardware's in a plain launch, the
// problem's own in a multi-problem launch (each problem keeps its grid, so its
// blocks compute, reduce and order exactly as in a launch of its own).
struct Blk {
  unsigned x, y, z, gx, gy, gz;
  __device__ __forceinline__ static Blk hw() {
    return Blk{blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x, gridDim.y, gridDim.z};
  }
  __device__ __forceinline__ unsigned lin() const { return x + gx * (y + gy * z); }
  __device__ __forceinline__ unsigned count() const { return gx * gy * gz; }
};

// Block partials -> last block -> fin(result[NV]) run by thread 0 of the last block.
//
// Ordering of a block's partials before its ticket.  The partials are
// agent-scope atomic stores, which write through the XCD's L2 (the `sc1`
// stores of gfx950); the block waits for their completion (s_waitcnt
// vmcnt(0)) and only then takes the ticket, and the last block reads the
// partials with agent-scope atomic loads.  This relies on the instruction
// order store -> wait -> ticket in the generated code, which a CPU test of
// the built assembly checks.  R.fence = 1 orders them with a release fence
// instead (and an acquire fence in the last block), as the memory model
// prescribes; on this GPU that fence writes back the XCD's whole L2, once
// per block.
template <int OP, int NV, class Fin>
__device__ __forceinline__ void reduce_finalize(const Blk& bk, double (&v)[NV], const Reduce& R, double init,
                                                Fin fin) {
  __shared__ double sh[NV * 16];
  __shared__ int is_last;
  block_red<OP, NV>(v, sh);
  const unsigned nb = bk.count();
  if (threadIdx.x == 0) {
    const unsigned b = bk.lin();
    for (int k = 0; k < NV; k++)
      __hip_atomic_store(&R.partials[(size_t)b * NV + k], v[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (R.fence)
      __threadfence();
    else
      __builtin_amdgcn_s_waitcnt(0);
    unsigned t = atomicAdd(R.counter, 1u);
    is_last = (t == nb - 1);
  }
  __syncthreads();
  if (!is_last) return;
  if (R.fence) __threadfence();
  // the partial loads below are agent-scope atomic loads, issued after the
  // last ticket: every block's partials completed before its ticket
  double acc[NV];
  for (int k = 0; k < NV; k++) acc[k] = init;
  for (unsigned i = threadIdx.x; i < nb; i += blockDim.x)
    for (int k = 0; k < NV; k++)
      acc[k] = red2<OP>(acc[k], __hip_atomic_load(&R.partials[(size_t)i * NV + k], __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT));
  block_red<OP, NV>(acc, sh);
  if (threadIdx.x == 0) {
    fin(acc);
    __hip_atomic_store(R.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}
template <int OP, int NV, class Fin>
__device__ __forceinline__ void reduce_finalize(double (&v)[NV], const Reduce& R, double init, Fin fin) {
  reduce_finalize<OP, NV>(Blk::hw(), v, R, init, fin);
}

// ---------------------------------------------------------------------------
// Element algebra (sum factorisation)
// ---------------------------------------------------------------------------
// Interpolation of a node pair (a, b) to the two Gauss points:
// (GA a + GB b, GB a + GA b).  With GA + GB = 1 this is b + GA (a - b) and
// a - GA (a - b): three operations instead of four (the "delta form"), used
// by both precisions.  HPCR_DELTA_INTERP_F64=0 keeps the two-product form in
// FP64 (about 7 % slower fine-level passes, the same operator up to rounding).
#ifndef HPCR_DELTA_INTERP_F64
#define HPCR_DELTA_INTERP_F64 1
#endif
template <class T>
constexpr bool kDelta = !std::is_same<T, double>::value || HPCR_DELTA_INTERP_F64;

template <class T>
__device__ __forceinline__ void lerp2(T a, T b, T& o0, T& o1) {
  if constexpr (kDelta<T>) {
    const T d = a - b;
    o0 = fma(T(GA), d, b);
    o1 = fma(T(-GA), d, a);
  } else {
    o0 = T(GA) * a + T(GB) * b;
    o1 = T(GB) * a + T(GA) * b;
  }
}

// trilinear interpolation of 8 nodal values v[iz][iy][ix] to Gauss points o[gz][gy][gx]
template <class T>
__device__ __forceinline__ void interp8(const T (&v)[2][2][2], T (&o)[2][2][2]) {
  T t[2][2][2], s[2][2][2];
#pragma unroll
  for (int iz = 0; iz < 2; iz++)
#pragma unroll
    for (int iy = 0; iy < 2; iy++) lerp2(v[iz][iy][0], v[iz][iy][1], t[iz][iy][0], t[iz][iy][1]);
#pragma unroll
  for (int iz = 0; iz < 2; iz++)
#pragma unroll
    for (int gx = 0; gx < 2; gx++) lerp2(t[iz][0][gx], t[iz][1][gx], s[iz][0][gx], s[iz][1][gx]);
#pragma unroll
  for (int gy = 0; gy < 2; gy++)
#pragma unroll
    for (int gx = 0; gx < 2; gx++) lerp2(s[0][gy][gx], s[1][gy][gx], o[0][gy][gx], o[1][gy][gx]);
}
// interp8 split at the node planes, for values carried up a z-march: the x-
// and y-passes of one node plane p[iy][ix] (s[gy][gx]), and the z-pass
// between the bottom and top planes' results.  interp8 runs the same
// operations in the same order.
template <class T>
__device__ __forceinline__ void interp_plane(const T (&p)[2][2], T (&s)[2][2]) {
  T t[2][2];
#pragma unroll
  for (int iy = 0; iy < 2; iy++) lerp2(p[iy][0], p[iy][1], t[iy][0], t[iy][1]);
#pragma unroll
  for (int gx = 0; gx < 2; gx++) lerp2(t[0][gx], t[1][gx], s[0][gx], s[1][gx]);
}
template <class T>
__device__ __forceinline__ void interp_z(const T (&s0)[2][2], const T (&s1)[2][2], T (&o)[2][2][2]) {
#pragma unroll
  for (int gy = 0; gy < 2; gy++)
#pragma unroll
    for (int gx = 0; gx < 2; gx++) lerp2(s0[gy][gx], s1[gy][gx], o[0][gy][gx], o[1][gy][gx]);
}

// transpose of interp8: nodal f[iz][iy][ix] (+)= sum_g L L L q[gz][gy][gx]
// (SET assigns instead of adding to f)
template <bool SET = false, class T>
__device__ __forceinline__ void interp8_t_add(const T (&q)[2][2][2], T (&f)[2][2][2]) {
  T s[2][2][2], t[2][2][2];
#pragma unroll
  for (int gy = 0; gy < 2; gy++)
#pragma unroll
    for (int gx = 0; gx < 2; gx++) lerp2(q[0][gy][gx], q[1][gy][gx], s[0][gy][gx], s[1][gy][gx]);
#pragma unroll
  for (int iz = 0; iz < 2; iz++)
#pragma unroll
    for (int gx = 0; gx < 2; gx++) lerp2(s[iz][0][gx], s[iz][1][gx], t[iz][0][gx], t[iz][1][gx]);
#pragma unroll
  for (int iz = 0; iz < 2; iz++)
#pragma unroll
    for (int iy = 0; iy < 2; iy++) {
      T a, b;
      lerp2(t[iz][iy][0], t[iz][iy][1], a, b);
      if constexpr (SET) {
        f[iz][iy][0] = a;
        f[iz][iy][1] = b;
      } else {
        f[iz][iy][0] += a;
        f[iz][iy][1] += b;
      }
    }
}
// bilinear interpolation of 4 edge values e[a][b] (a: outer axis, b: inner
// axis) to Gauss pairs o[ga][gb], scaled by s (the delta form scales the
// result, the two-product form the inner stage).  SC false: no scaling (the
// caller has folded the grid scale into the moduli).
template <bool SC = true, class T>
__device__ __forceinline__ void interp4(const T (&e)[2][2], T s, T (&o)[2][2]) {
  if constexpr (kDelta<T>) {
    T t0, t1, u0, u1;
    lerp2(e[0][0], e[0][1], t0, t1);
    lerp2(e[1][0], e[1][1], u0, u1);
    lerp2(t0, u0, o[0][0], o[1][0]);
    lerp2(t1, u1, o[0][1], o[1][1]);
    if constexpr (SC) {
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) o[a][b] *= s;
    }
  } else {
    const T ga = T(GA), gb = T(GB);
    auto sc = [s](T v) { return SC ? s * v : v; };
    const T t0 = sc(ga * e[0][0] + gb * e[0][1]);
    const T t1 = sc(gb * e[0][0] + ga * e[0][1]);
    const T u0 = sc(ga * e[1][0] + gb * e[1][1]);
    const T u1 = sc(gb * e[1][0] + ga * e[1][1]);
    o[0][0] = ga * t0 + gb * u0;
    o[0][1] = ga * t1 + gb * u1;
    o[1][0] = gb * t0 + ga * u0;
    o[1][1] = gb * t1 + ga * u1;
  }
}
// transpose of interp4
template <bool SC = true, class T>
__device__ __forceinline__ void interp4_t(const T (&S)[2][2], T s, T (&R)[2][2]) {
  if constexpr (kDelta<T>) {
    T q00, q01, q10, q11;
    lerp2(S[0][0], S[1][0], q00, q10);
    lerp2(S[0][1], S[1][1], q01, q11);
    lerp2(q00, q01, R[0][0], R[0][1]);
    lerp2(q10, q11, R[1][0], R[1][1]);
    if constexpr (SC) {
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) R[a][b] *= s;
    }
  } else {
    const T ga = T(GA), gb = T(GB);
    auto sc = [s](T v) { return SC ? s * v : v; };
    const T q00 = ga * S[0][0] + gb * S[1][0];  // a=0, gb=0
    const T q01 = ga * S[0][1] + gb * S[1][1];
    const T q10 = gb * S[0][0] + ga * S[1][0];  // a=1
    const T q11 = gb * S[0][1] + ga * S[1][1];
    R[0][0] = sc(ga * q00 + gb * q01);
    R[0][1] = sc(gb * q00 + ga * q01);
    R[1][0] = sc(ga * q10 + gb * q11);
    R[1][1] = sc(gb * q10 + ga * q11);
  }
}

// Gradient of a nodal scalar v[iz][iy][ix] at the Gauss points:
// gxv[gz][gy] (d/dx), gyv[gz][gx] (d/dy), gzv[gy][gx] (d/dz)
template <bool SC = true, class T>
__device__ __forceinline__ void grad8(const T (&v)[2][2][2], T sx, T sy, T sz, T (&gxv)[2][2], T (&gyv)[2][2],
                                      T (&gzv)[2][2]) {
  T ex[2][2], ey[2][2], ez[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      ex[a][b] = v[a][b][1] - v[a][b][0];  // [iz][iy]
      ey[a][b] = v[a][1][b] - v[a][0][b];  // [iz][ix]
      ez[a][b] = v[1][a][b] - v[0][a][b];  // [iy][ix]
    }
  interp4<SC>(ex, sx, gxv);
  interp4<SC>(ey, sy, gyv);
  interp4<SC>(ez, sz, gzv);
}
// grad8 split at the node planes: the in-plane parts of d/dx and d/dy of one
// plane p[iy][ix] (its x-edge differences interpolated over y, its y-edge
// differences over x), and the gradient from the bottom and top planes' parts
// and the element's z-edges.  grad8 runs the same operations in the same
// order.
template <class T>
__device__ __forceinline__ void grad_plane(const T (&p)[2][2], T (&xp)[2], T (&yp)[2]) {
  lerp2(p[0][1] - p[0][0], p[1][1] - p[1][0], xp[0], xp[1]);
  lerp2(p[1][0] - p[0][0], p[1][1] - p[0][1], yp[0], yp[1]);
}
template <bool SC = true, class T>
__device__ __forceinline__ void grad8_planes(const T (&v)[2][2][2], const T (&xb)[2], const T (&yb)[2],
                                             const T (&xt)[2], const T (&yt)[2], T sx, T sy, T sz,
                                             T (&gxv)[2][2], T (&gyv)[2][2], T (&gzv)[2][2]) {
  lerp2(xb[0], xt[0], gxv[0][0], gxv[1][0]);
  lerp2(xb[1], xt[1], gxv[0][1], gxv[1][1]);
  lerp2(yb[0], yt[0], gyv[0][0], gyv[1][0]);
  lerp2(yb[1], yt[1], gyv[0][1], gyv[1][1]);
  T ez[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) ez[a][b] = v[1][a][b] - v[0][a][b];
  interp4<SC>(ez, sz, gzv);
  if constexpr (SC) {
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++) {
        gxv[a][b] *= sx;
        gyv[a][b] *= sy;
      }
  }
}
// Transposed gradient: f[iz][iy][ix] = sum_g dN/dx_j * S_j (SET) or
// f += (SET false).  Assigning the x part instead of adding it to 0 saves
// the additions of zero, which the compiler keeps (0 + x is +0 for x = -0);
// every other value is the same.
template <bool SET, bool SC = true, class T>
__device__ __forceinline__ void grad8_t(const T (&Sx)[2][2], const T (&Sy)[2][2], const T (&Sz)[2][2], T sx, T sy,
                                        T sz, T (&f)[2][2][2]) {
  T R[2][2];
  interp4_t<SC>(Sx, sx, R);  // R[iz][iy]
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      if constexpr (SET) {
        f[a][b][1] = R[a][b];
        f[a][b][0] = -R[a][b];
      } else {
        f[a][b][1] += R[a][b];
        f[a][b][0] -= R[a][b];
      }
    }
  interp4_t<SC>(Sy, sy, R);  // R[iz][ix]
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      f[a][1][b] += R[a][b];
      f[a][0][b] -= R[a][b];
    }
  interp4_t<SC>(Sz, sz, R);  // R[iy][ix]
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      f[1][a][b] += R[a][b];
      f[0][a][b] -= R[a][b];
    }
}
template <class T>
__device__ __forceinline__ void grad8_t_add(const T (&Sx)[2][2], const T (&Sy)[2][2], const T (&Sz)[2][2], T sx,
                                            T sy, T sz, T (&f)[2][2][2]) {
  grad8_t<false>(Sx, Sy, Sz, sx, sy, sz, f);
}

// true when any of the 8 values is 0 (+0 or -0)
template <class T>
__device__ __forceinline__ bool any_zero(const T (&v)[2][2][2]) {
  bool z = false;
#pragma unroll
  for (int i = 0; i < 8; i++) z |= (v[i >> 2][(i >> 1) & 1][i & 1] == T(0));
  return z;
}

// a product the compiler may not fuse into a neighbouring addition
template <class T>
__device__ __forceinline__ T mul_nc(T a, T b) {
#pragma clang fp contract(off)
  return a * b;
}

// gauss_point_MP (material_gauss_point.cpp:18-41): zero if any nodal value is 0
template <class T>
__device__ __forceinline__ void gp_mp(const T (&v)[2][2][2], T (&o)[2][2][2]) {
  const bool z = any_zero(v);
  interp8(v, o);
  if (z) {
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int c = 0; c < 2; c++) o[a][b][c] = T(0);
  }
}
// (1-k)(1-d)^2 + k with d clamped to [0,1] (material_gauss_point.cpp:50-56).
// The clamp is a max and a min: for every d that is not a NaN it equals the
// reference's two compares (d < 0 -> 0, d > 1 -> 1; a d of -0 squares to the
// same factor), in two instructions where the compares and selects take six.
template <class T>
__device__ __forceinline__ T degr(T dg, T k) {
  dg = fmin(fmax(dg, T(0)), T(1));
  return (T(1) - k) * (T(1) - dg) * (T(1) - dg) + k;
}

// Gauss-point Lame coefficients (lambda = K_d - 2/3 G_d, mu = G_d) of a fine
// element (MP_all_GAUSS_Point, mechanical_pcg.cpp:26-61).
// K_d = (aK + bK g) K_g and G_d = (aG + bG g) G_g with a, b in {0, 1}, so the
// factors are exactly g, 1 or 0: the split flag picks which modulus degrades
// by g = degr(d_g), and the zero rule (a nodal K or G of 0) makes the modulus 0
// at every Gauss point.  The products are kept out of fused multiply-adds, so
// the values equal the reference's selected ones.  An element with a node at
// d == 0 has d = 0 at every Gauss point (the zero rule on d), i.e. the
// undamaged factor degr(0, k) everywhere: that common case skips the
// interpolation of d.
template <class T>
__device__ __forceinline__ void fine_lame(const T (&Kn)[2][2][2], const T (&Gn)[2][2][2], const T (&dn)[2][2][2],
                                          int flag, T small_k, T (&lam)[2][2][2], T (&mu)[2][2][2]) {
  const bool zK = any_zero(Kn), zG = any_zero(Gn);
  const T bK = (!zK && flag == 0) ? T(1) : T(0), aK = (!zK && flag != 0) ? T(1) : T(0);
  const T bG = (!zG && flag != 2) ? T(1) : T(0), aG = (!zG && flag == 2) ? T(1) : T(0);
  T Kg[2][2][2], Gg[2][2][2];
  interp8(Kn, Kg);
  interp8(Gn, Gg);
  if (any_zero(dn)) {
    const T g0 = degr(T(0), small_k);
    const T gK = aK + bK * g0, gG = aG + bG * g0;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int c = 0; c < 2; c++) {
          const T Gd = mul_nc(gG, Gg[a][b][c]);
          lam[a][b][c] = mul_nc(gK, Kg[a][b][c]) - (T(2) / T(3)) * Gd;
          mu[a][b][c] = Gd;
        }
  } else {
    T dg[2][2][2];
    interp8(dn, dg);
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int c = 0; c < 2; c++) {
          const T g = degr(dg[a][b][c], small_k);
          const T Gd = mul_nc(aG + bG * g, Gg[a][b][c]);
          lam[a][b][c] = mul_nc(aK + bK * g, Kg[a][b][c]) - (T(2) / T(3)) * Gd;
          mu[a][b][c] = Gd;
        }
  }
}

// What a fine element passes up the z-march to the element above it: its top
// node plane's x/y-interpolated K and G, whether that plane has a zero K, G or
// d node, and its in-plane gradient parts (grad_plane) of u.
template <class T>
struct FineCarry {
  T Ks[2][2], Gs[2][2];
  bool zK, zG, zd;
  T gx[3][2], gy[3][2];
};
template <class T>
__device__ __forceinline__ bool any_zero4(const T (&p)[2][2]) {
  return (p[0][0] == T(0)) | (p[0][1] == T(0)) | (p[1][0] == T(0)) | (p[1][1] == T(0));
}

// fine_lame with the bottom plane's part of the work taken from the element
// below (first: the lowest element of a block's march computes both planes);
// the same values as fine_lame
template <class T>
__device__ __forceinline__ void fine_lame_planes(const T (&Kn)[2][2][2], const T (&Gn)[2][2][2],
                                                 const T (&dn)[2][2][2], int flag, T small_k, FineCarry<T>& cr,
                                                 bool first, T (&lam)[2][2][2], T (&mu)[2][2][2]) {
  T Kb[2][2], Gb[2][2];
  bool zKb, zGb, zdb;
  if (first) {
    interp_plane(Kn[0], Kb);
    interp_plane(Gn[0], Gb);
    zKb = any_zero4(Kn[0]);
    zGb = any_zero4(Gn[0]);
    zdb = any_zero4(dn[0]);
  } else {
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++) {
        Kb[a][b] = cr.Ks[a][b];
        Gb[a][b] = cr.Gs[a][b];
      }
    zKb = cr.zK;
    zGb = cr.zG;
    zdb = cr.zd;
  }
  interp_plane(Kn[1], cr.Ks);
  interp_plane(Gn[1], cr.Gs);
  cr.zK = any_zero4(Kn[1]);
  cr.zG = any_zero4(Gn[1]);
  cr.zd = any_zero4(dn[1]);
  const bool zK = zKb | cr.zK, zG = zGb | cr.zG;
  const T bK = (!zK && flag == 0) ? T(1) : T(0), aK = (!zK && flag != 0) ? T(1) : T(0);
  const T bG = (!zG && flag != 2) ? T(1) : T(0), aG = (!zG && flag == 2) ? T(1) : T(0);
  T Kg[2][2][2], Gg[2][2][2];
  interp_z(Kb, cr.Ks, Kg);
  interp_z(Gb, cr.Gs, Gg);
  if (zdb | cr.zd) {
    const T g0 = degr(T(0), small_k);
    const T gK = aK + bK * g0, gG = aG + bG * g0;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int c = 0; c < 2; c++) {
          const T Gd = mul_nc(gG, Gg[a][b][c]);
          lam[a][b][c] = mul_nc(gK, Kg[a][b][c]) - (T(2) / T(3)) * Gd;
          mu[a][b][c] = Gd;
        }
  } else {
    T dg[2][2][2];
    interp8(dn, dg);
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int c = 0; c < 2; c++) {
          const T g = degr(dg[a][b][c], small_k);
          const T Gd = mul_nc(aG + bG * g, Gg[a][b][c]);
          lam[a][b][c] = mul_nc(aK + bK * g, Kg[a][b][c]) - (T(2) / T(3)) * Gd;
          mu[a][b][c] = Gd;
        }
  }
}

// The element's forces from its displacement gradients at the Gauss points:
// stresses, then the transposed contraction (the second half of
// mech_element).
template <bool CONST_COEF, bool SC = true, class T>
__device__ __forceinline__ void mech_element_from_grad(const T (&dx)[3][2][2], const T (&dy)[3][2][2],
                                                       const T (&dz)[3][2][2], const T (&lam)[2][2][2],
                                                       const T (&mu)[2][2][2], T lam0, T mu0, T tx, T ty, T tz,
                                                       T (&f)[3][2][2][2]) {
  // Sx sums over gx, Sy over gy, Sz over gz: the first Gauss point of each
  // sum assigns, the second adds (no additions of zero)
  T Sx[3][2][2], Sy[3][2][2], Sz[3][2][2];
  auto put = [](T& s, T v, bool first) {
    if (first) s = v; else s += v;
  };
#pragma unroll
  for (int gz = 0; gz < 2; gz++)
#pragma unroll
    for (int gy = 0; gy < 2; gy++)
#pragma unroll
      for (int gx = 0; gx < 2; gx++) {
        const T exx = dx[0][gz][gy], eyy = dy[1][gz][gx], ezz = dz[2][gy][gx];
        const T gxy = dx[1][gz][gy] + dy[0][gz][gx];
        const T gxz = dx[2][gz][gy] + dz[0][gy][gx];
        const T gyz = dy[2][gz][gx] + dz[1][gy][gx];
        const T tr = exx + eyy + ezz;
        const T l = CONST_COEF ? lam0 : lam[gz][gy][gx];
        const T m = CONST_COEF ? mu0 : mu[gz][gy][gx];
        const T lt = l * tr, m2 = T(2) * m;
        const T sxx = lt + m2 * exx, syy = lt + m2 * eyy, szz = lt + m2 * ezz;
        const T sxy = m * gxy, sxz = m * gxz, syz = m * gyz;
        put(Sx[0][gz][gy], sxx, gx == 0); put(Sx[1][gz][gy], sxy, gx == 0); put(Sx[2][gz][gy], sxz, gx == 0);
        put(Sy[0][gz][gx], sxy, gy == 0); put(Sy[1][gz][gx], syy, gy == 0); put(Sy[2][gz][gx], syz, gy == 0);
        put(Sz[0][gy][gx], sxz, gz == 0); put(Sz[1][gy][gx], syz, gz == 0); put(Sz[2][gy][gx], szz, gz == 0);
      }
#pragma unroll
  for (int c = 0; c < 3; c++) grad8_t<true, SC>(Sx[c], Sy[c], Sz[c], tx, ty, tz, f[c]);
}

// y_e = K_e u_e for one element: u[c][iz][iy][ix] -> f[c][iz][iy][ix]
// (Internal_Force_element, mechanical_pcg.cpp:114-172).  SC false: the
// gradient and its transpose are left unscaled (sx..tz unused), for callers
// that have folded the grid scale into the moduli: on a grid with
// hx = hy = hz = h, every term of f carries the factor (1/h)(w_g/h) =
// w_g/h^2, and lambda, mu enter f linearly.
template <bool CONST_COEF, bool SC = true, class T>
__device__ __forceinline__ void mech_element(const T (&u)[3][2][2][2], const T (&lam)[2][2][2],
                                             const T (&mu)[2][2][2], T lam0, T mu0, T sx, T sy, T sz, T tx, T ty,
                                             T tz, T (&f)[3][2][2][2]) {
  T dx[3][2][2], dy[3][2][2], dz[3][2][2];
#pragma unroll
  for (int c = 0; c < 3; c++) grad8<SC>(u[c], sx, sy, sz, dx[c], dy[c], dz[c]);
  mech_element_from_grad<CONST_COEF, SC>(dx, dy, dz, lam, mu, lam0, mu0, tx, ty, tz, f);
}

// mech_element<false, SC> with the gradient's bottom-plane parts taken from the
// element below (grad8_planes); the same values as mech_element
template <bool SC = true, class T>
__device__ __forceinline__ void mech_element_planes(const T (&u)[3][2][2][2], const T (&lam)[2][2][2],
                                                    const T (&mu)[2][2][2], FineCarry<T>& cr, bool first, T sx,
                                                    T sy, T sz, T tx, T ty, T tz, T (&f)[3][2][2][2]) {
  static_assert(kDelta<T>, "the plane-split gradient follows the delta-form interpolation");
  T dx[3][2][2], dy[3][2][2], dz[3][2][2];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    T xb[2], yb[2];
    if (first) {
      grad_plane(u[c][0], xb, yb);
    } else {
      xb[0] = cr.gx[c][0]; xb[1] = cr.gx[c][1];
      yb[0] = cr.gy[c][0]; yb[1] = cr.gy[c][1];
    }
    grad_plane(u[c][1], cr.gx[c], cr.gy[c]);
    grad8_planes<SC>(u[c], xb, yb, cr.gx[c], cr.gy[c], sx, sy, sz, dx[c], dy[c], dz[c]);
  }
  mech_element_from_grad<false, SC>(dx, dy, dz, lam, mu, T(0), T(0), tx, ty, tz, f);
}

// Jacobi diagonal of one element (PreM_u_element, mechanical_pcg.cpp:63-100)
__device__ __forceinline__ void mech_element_diag(const double (&lam)[2][2][2], const double (&mu)[2][2][2],
                                                  double ihx2, double ihy2, double ihz2, double wg,
                                                  double (&f)[3][2][2][2]) {
#pragma unroll
  for (int ka = 0; ka < 2; ka++)
#pragma unroll
    for (int ja = 0; ja < 2; ja++)
#pragma unroll
      for (int ia = 0; ia < 2; ia++) {
        double s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
        for (int gz = 0; gz < 2; gz++)
#pragma unroll
          for (int gy = 0; gy < 2; gy++)
#pragma unroll
            for (int gx = 0; gx < 2; gx++) {
              const double lx = Lw(ia, gx), ly = Lw(ja, gy), lz = Lw(ka, gz);
              const double X = ihx2 * ly * ly * lz * lz;
              const double Y = ihy2 * lx * lx * lz * lz;
              const double Z = ihz2 * lx * lx * ly * ly;
              const double l = lam[gz][gy][gx], m = mu[gz][gy][gx];
              const double p = l + 2.0 * m;
              s0 += p * X + m * (Y + Z);
              s1 += p * Y + m * (X + Z);
              s2 += p * Z + m * (X + Y);
            }
        f[0][ka][ja][ia] = wg * s0;
        f[1][ka][ja][ia] = wg * s1;
        f[2][ka][ja][ia] = wg * s2;
      }
}

// x / b for a kernel-uniform b, with rb = 1/b rounded once on the host: the
// product x rb corrected by one fused remainder step (Markstein), which gives
// the correctly rounded quotient -- three instructions instead of the full
// division sequence.  Checked bit for bit against IEEE division on 3e8
// random operands over 15 divisors.
__device__ __forceinline__ double div_uniform(double x, double b, double rb) {
  const double q = mul_nc(x, rb);
  return fma(fma(-q, b, x), rb, q);
}

// Phase-field element (Fld_element, phase_field_pcg.cpp:61-91); rlc = 1/lc
__device__ __forceinline__ void pf_coefs(const double (&gcn)[2][2][2], double He, double lc, double rlc, double wg,
                                         double (&m)[2][2][2], double (&k)[2][2][2]) {
  double gg[2][2][2];
  gp_mp(gcn, gg);
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int c = 0; c < 2; c++) {
        m[a][b][c] = wg * (2.0 * He + div_uniform(gg[a][b][c], lc, rlc));
        k[a][b][c] = wg * gg[a][b][c] * lc;
      }
}
__device__ __forceinline__ void pf_element(const double (&v)[2][2][2], const double (&m)[2][2][2],
                                           const double (&k)[2][2][2], double sx, double sy, double sz,
                                           double (&f)[2][2][2]) {
  double dv[2][2][2];
  interp8(v, dv);
  double q[2][2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int c = 0; c < 2; c++) q[a][b][c] = m[a][b][c] * dv[a][b][c];
  interp8_t_add<true>(q, f);
  double gxv[2][2], gyv[2][2], gzv[2][2];
  grad8(v, sx, sy, sz, gxv, gyv, gzv);
  double Sx[2][2], Sy[2][2], Sz[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      Sx[a][b] = gxv[a][b] * (k[a][b][0] + k[a][b][1]);  // [gz][gy], sum over gx
      Sy[a][b] = gyv[a][b] * (k[a][0][b] + k[a][1][b]);  // [gz][gx], sum over gy
      Sz[a][b] = gzv[a][b] * (k[0][a][b] + k[1][a][b]);  // [gy][gx], sum over gz
    }
  grad8_t_add(Sx, Sy, Sz, sx, sy, sz, f);
}
// Md_element (phase_field_pcg.cpp:20-47)
__device__ __forceinline__ void pf_element_diag(const double (&m)[2][2][2], const double (&k)[2][2][2],
                                                double ihx2, double ihy2, double ihz2, double (&f)[2][2][2]) {
#pragma unroll
  for (int ka = 0; ka < 2; ka++)
#pragma unroll
    for (int ja = 0; ja < 2; ja++)
#pragma unroll
      for (int ia = 0; ia < 2; ia++) {
        double s = 0;
#pragma unroll
        for (int gz = 0; gz < 2; gz++)
#pragma unroll
          for (int gy = 0; gy < 2; gy++)
#pragma unroll
            for (int gx = 0; gx < 2; gx++) {
              const double lx = Lw(ia, gx), ly = Lw(ja, gy), lz = Lw(ka, gz);
              const double N2 = lx * lx * ly * ly * lz * lz;
              const double G2 = ihx2 * ly * ly * lz * lz + ihy2 * lx * lx * lz * lz + ihz2 * lx * lx * ly * ly;
              s += m[gz][gy][gx] * N2 + k[gz][gy][gx] * G2;
            }
        f[ka][ja][ia] = s;
      }
}

// Dirichlet DOF test of the reference's boundary conditions
// (boundary_conditions.cpp:28-66): returns a mask, bit c set when component
// c is prescribed; bit 3 marks u_z = Ut (top face) rather than 0.
// From the node's three axis flags:
__device__ __forceinline__ int dir_mask_bits(unsigned fz, unsigned fx, unsigned fy) {
  int m = 0;
  if ((fz & 1) && (fx & 1) && (fy & 1)) m |= 1;
  if ((fz & 1) && (fx & 2) && (fy & 2)) m |= 2;
  if (fz & 1) m |= 4;
  if (fz & 2) m |= 4 | 8;
  return m;
}
__device__ __forceinline__ int dir_mask(const LevelGeom& L, int i, int j, int k) {
  const unsigned fz = L.fz[k];
  if (fz == 0) return 0;
  return dir_mask_bits(fz, L.fx[i], L.fy[j]);
}

// The axis flags of a level staged in LDS (nx + ny + nz bytes), so that a
// node's mask costs LDS reads rather than dependent global loads.  All
// threads of the block call stage(); the caller synchronises.
constexpr int DIRF_MAX = 12288;
struct DirFlags {
  unsigned char* s;
  __device__ __forceinline__ void stage(const LevelGeom& L) const {
    const int n = L.nx + L.ny + L.nz;
    for (int t = threadIdx.x; t < n; t += blockDim.x)
      s[t] = t < L.nx ? L.fx[t] : (t < L.nx + L.ny ? L.fy[t - L.nx] : L.fz[t - L.nx - L.ny]);
  }
  __device__ __forceinline__ int mask(const LevelGeom& L, int i, int j, int k) const {
    const unsigned fz = s[L.nx + L.ny + k];
    return fz ? dir_mask_bits(fz, s[i], s[L.nx + j]) : 0;
  }
};

// ---------------------------------------------------------------------------
// Tile kernel framework
// ---------------------------------------------------------------------------
// lane i <- lane i-1 across the wavefront (DPP wave_shr:1); lane 0 gets 0
// (bound_ctrl, so no old value has to be materialised).  Element (tx, ty) and
// its x-neighbour (tx-1, ty) are adjacent lanes because TX divides the
// wavefront into whole rows; lanes with tx == 0 own no node.
__device__ __forceinline__ float wave_shr1(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ double wave_shr1(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), 0x138, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x138, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous range of tiles so that tiles sharing halo rows share an L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int q = nb / 8, rem = nb % 8;
  const int xcd = b % 8, j = b / 8;
  return (xcd < rem ? xcd * (q + 1) : rem * (q + 1) + (xcd - rem) * q) + j;
}

// optional occupancy target of the tile kernels (waves per SIMD), a tuning macro
#ifdef HPCR_TILE_WAVES_PER_EU
#define HPCR_TILE_ATTR __attribute__((amdgpu_waves_per_eu(HPCR_TILE_WAVES_PER_EU)))
#else
#define HPCR_TILE_ATTR
#endif

// An op whose element passes values up the z-march declares `Carry` (not
// void) and element_planes(..., Carry&, bool first).
template <class Op, class = void>
struct CarryOf {
  using type = char;
  static constexpr bool value = false;
};
template <class Op>
struct CarryOf<Op, std::void_t<typename Op::Carry>> {
  using type = std::conditional_t<std::is_void<typename Op::Carry>::value, char, typename Op::Carry>;
  static constexpr bool value = !std::is_void<typename Op::Carry>::value;
};

template <class Op>
__device__ __forceinline__ void tile_body(const LevelGeom& L, const Op& op, const Blk& bk) {
  constexpr int NC = Op::NC;  // output components
  constexpr int NV = Op::NV;  // values per node in LDS
  constexpr int NR = Op::NR;  // raw values per node fetched from HBM
  using T = typename Op::T;   // element arithmetic (double, or float in the preconditioner)
  using TR = typename Op::TR; // type of the fetched raw values
  constexpr int NE = Op::NE;  // per-element values (split flag, K_e/G_e, H, Gauss-point moduli)
  using TE = typename Op::TE;
  static_assert(NE >= 1, "tile ops carry per-element data");
  constexpr int NS = Op::NSL;  // node-plane slots in LDS (2 or 3)
  static_assert(NS == 2 || NS == 3, "two or three plane slots");
  __shared__ T sN[NS][NV][PLANE];
  // per-element data travels with the node planes: layer p sits in slot(p)
  __shared__ TE sE[NS][NE][TPB];
  // y-exchange buffers; with three plane slots they alternate between layers
  constexpr int NXB = (NS == 3) ? 2 : 1;
  __shared__ T sX[NXB][2 * NC][TPB];
  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  const int ntx = (L.nx + OWNX - 1) / OWNX, nty = (L.ny + OWNY - 1) / OWNY;
  const int t = xcd_tile(bk.x, bk.gx);
  const int bx0 = (t % ntx) * OWNX, by0 = ((t / ntx) % nty) * OWNY;
  const int z0 = (t / (ntx * nty)) * L.zc;
  const int zend = min(z0 + L.zc, L.nz);
  const int ex = bx0 - 1 + tx, ey = by0 - 1 + ty;
  const bool owner = (tx >= 1) && (ty >= 1) && (ex < L.nx) && (ey < L.ny);
  const bool exy = (ex >= 0) && (ey >= 0) && (ex <= L.nx - 2) && (ey <= L.ny - 2);
  const typename Op::State st = op.state();

  // node slots of this thread in a plane: tid and tid + TPB
  const bool has2 = tid + TPB < PLANE;
  const int q1 = tid, q2 = tid + TPB;
  const int gx1 = bx0 - 1 + q1 % PX, gy1 = by0 - 1 + q1 / PX;
  const int gx2 = bx0 - 1 + q2 % PX, gy2 = by0 - 1 + q2 / PX;
  const bool in1 = gx1 >= 0 && gy1 >= 0 && gx1 < L.nx && gy1 < L.ny;
  const bool in2 = has2 && gx2 >= 0 && gy2 >= 0 && gx2 < L.nx && gy2 < L.ny;
  typename Op::Raw r1, r2;  // raw node values in flight (Op::Raw is TR[NR] or an op's own record)
  TE re[NE];
  // node plane pz and the element layer above it (planes past zend are not needed)
  auto fetch = [&](int pz) {
    const bool zok = pz >= 0 && pz < L.nz && pz <= zend;
    op.efetch(L, ex, ey, pz, exy && zok && pz < zend && pz <= L.nz - 2, re);
    op.fetch(L, gx1, gy1, pz, in1 && zok, st, r1);
    op.fetch(L, gx2, gy2, pz, in2 && zok, st, r2);
  };
  auto commit = [&](int slot) {
#pragma unroll
    for (int k = 0; k < NE; k++) sE[slot][k][tid] = re[k];
    T v[NV];
    op.commit(r1, st, v);
#pragma unroll
    for (int k = 0; k < NV; k++) sN[slot][k][q1] = v[k];
    if (has2) {
      op.commit(r2, st, v);
#pragma unroll
      for (int k = 0; k < NV; k++) sN[slot][k][q2] = v[k];
    }
  };

  // Dirichlet flags of the owner node: x and y once, the chunk's z flags in
  // LDS (visible after the prologue's barrier)
  __shared__ unsigned char sFz[ZCHUNK_MAX];
  if (tid < zend - z0) sFz[tid] = L.fz[z0 + tid];
  const unsigned fxo = owner ? L.fx[ex] : 0u, fyo = owner ? L.fy[ey] : 0u;

  // LDS slot of node plane p >= -1
  auto slot = [](int p) { return (NS == 3) ? (p + 3) % 3 : (p & 1); };

  // Inputs of the epilogue at this thread's own node, read one plane ahead.
  // The load is unconditional, from a clamped (always valid) node, so the
  // registers carried into the next layer never need a merge copy that
  // would wait for the load; threads and planes without an epilogue ignore
  // what they read.
  typename Op::NData nd;
  const int exc = min(max(ex, 0), L.nx - 1), eyc = min(max(ey, 0), L.ny - 1);

  typename Op::Acc acc;
  typename CarryOf<Op>::type up{};  // what each element passes to the one above
  T carry[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) carry[c] = T(0);

  // planes z0-1 .. z0+NS-2 into LDS, the next one in flight
  fetch(z0 - 1);
  commit(slot(z0 - 1));
  fetch(z0);
  commit(slot(z0));
  fetch(z0 + 1);
  if constexpr (NS == 3) {
    commit(slot(z0 + 1));
    fetch(z0 + 2);
  }
  __syncthreads();
  // One layer of the march.  The layer below the chunk (kz = z0 - 1) runs
  // first and on its own, so that an op with a carry gets `first` as a
  // compile-time constant.
  auto layer = [&](int kz, auto first) {
    const int s0 = slot(kz);
    const int xb = (NXB == 2) ? (kz & 1) : 0;
    // Every thread computes its element.  One outside the grid (a node off
    // the grid, or layer -1 / nz-1) finds zeros in LDS for the off-grid nodes
    // and for its element data, which zeroes all of its coefficients (the
    // zero rule of gauss_point_MP on K, G and gc; K_e = G_e = H = 0 on coarse
    // levels), so its contributions are exactly 0 without a branch.
    T f[NC][2][2][2];
    {
      TE e[NE];
#pragma unroll
      for (int k = 0; k < NE; k++) e[k] = sE[s0][k][tid];
      if constexpr (CarryOf<Op>::value)
        op.element_planes(L, sN, s0, tx, ty, ex, ey, kz, st, e, f, up, decltype(first)::value);
      else
        op.element(L, sN, s0, tx, ty, ex, ey, kz, st, e, f);
    }
    // x-direction: node (tx, .) takes corner 0 of its own element and corner 1
    // of element tx-1 (adjacent lane); g[c][iz][iy]
    T g[NC][2][2];
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
      for (int iz = 0; iz < 2; iz++)
#pragma unroll
        for (int iy = 0; iy < 2; iy++) g[c][iz][iy] = f[c][iz][iy][0] + wave_shr1(f[c][iz][iy][1]);
    // y-direction through LDS: the iy = 1 halves go to the row above
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
      for (int iz = 0; iz < 2; iz++) sX[xb][c * 2 + iz][tid] = g[c][iz][1];
    // values of this thread's own node of plane kz, read before the slot is reused
    T own[NC];
    if constexpr (Op::NEEDS_OWN) {
#pragma unroll
      for (int c = 0; c < NC; c++) own[c] = sN[s0][c][tx + PX * ty];
    }
    __syncthreads();
    // plane kz is no longer read: its slot receives plane kz+NS and the
    // next plane goes in flight.  With three slots no thread reads this slot
    // before the next barrier, and the exchange buffer written in the next
    // layer is the other one, so that barrier is the only one per layer.
    commit(s0);
    fetch(kz + NS + 1);
    if (owner) {
      T vb[NC];
#pragma unroll
      for (int c = 0; c < NC; c++) {
        vb[c] = carry[c] + g[c][0][0] + sX[xb][c * 2 + 0][tid - TX];
        carry[c] = g[c][1][0] + sX[xb][c * 2 + 1][tid - TX];
      }
      if (kz >= z0) {
        const unsigned fz = sFz[kz - z0];
        const int dm = fz ? dir_mask_bits(fz, fxo, fyo) : 0;
        op.epilogue(L, own, ex, ey, kz, dm, vb, st, nd, acc);
      }
    }
    op.nfetch(L, exc, eyc, min(kz + 1, L.nz - 1), nd);
    if constexpr (NS == 2) __syncthreads();
    };
  layer(z0 - 1, std::true_type());
  for (int kz = z0; kz < zend; ++kz) layer(kz, std::false_type());
  op.finish(acc, bk);
}

template <class Op>
__global__ HPCR_TILE_ATTR __launch_bounds__(TPB, Op::MIN_WAVES) void tile_kernel(const LevelGeom L, const Op op) {
  tile_body<Op>(L, op, Blk::hw());
}

// row length of an LDS node plane of PL nodes: the 33x9 planes of the 32x8
// halo tiles, the 65x5 planes of the 64x4 split tiles
template <int PL>
__host__ __device__ constexpr int plane_row() {
  static_assert(PL == PLANE || PL == SPLANE, "a tile node plane");
  return PL == PLANE ? PX : SPX;
}

// the 8 corner values of element (tx, ty) from the LDS planes of its bottom
// (slot s0) and top (the next slot) nodes
template <class T, int NS, int NV, int PL>
__device__ __forceinline__ void gather8(const T (&sN)[NS][NV][PL], int s0, int tx, int ty, int k,
                                        T (&v)[2][2][2]) {
  constexpr int RX = plane_row<PL>();
  const int q = tx + RX * ty;
  const int s1 = (NS == 3) ? (s0 == 2 ? 0 : s0 + 1) : (s0 ^ 1);
#pragma unroll
  for (int iz = 0; iz < 2; iz++) {
    const int s = iz ? s1 : s0;
    v[iz][0][0] = sN[s][k][q];
    v[iz][0][1] = sN[s][k][q + 1];
    v[iz][1][0] = sN[s][k][q + RX];
    v[iz][1][1] = sN[s][k][q + RX + 1];
  }
}

// The z-march of the finest level's fused PCG pass A on split-output tiles
// (STX x STY elements; see SplitW).  Each element is computed once.  A wave
// is one 64-element row: the x-neighbour is the adjacent lane (DPP
// wave_shr:1, lane 0 receives 0: the element left of the tile is another
// tile's), the y-neighbour the next wave, through LDS.  Node (tx, ty) of the
// tile's region (x = ox + tx, y = oy + ty, tx < 64, ty < 4) takes the
// contributions of the tile's elements (tx-1..tx, ty-1..ty) and goes to w
// through the op's epilogue; the right column (tx = 64, from lane 63's x = 1
// corners), the top row (ty = 4, from row 3's iy = 1 halves) and their corner
// go to the side arrays through the op's side().  The side nodes' values of
// p and their z-carries stay in LDS (one writer and reader each), so that no
// registers are held for them across the layer's barrier.  A node's parts add
// up in pass B in a fixed order; the dot p.w is the sum of p times every part
// the block writes.
template <class Op>
__device__ __forceinline__ void tile_body_split(const LevelGeom& L, const Op& op, const SplitW& S, const Blk& bk) {
  constexpr int NC = Op::NC;
  constexpr int NV = Op::NV;
  using T = typename Op::T;
  constexpr int NE = Op::NE;
  using TE = typename Op::TE;
  static_assert(NC == 3 && Op::NEEDS_OWN && CarryOf<Op>::value, "the fused PCG pass A of the finest level");
  constexpr int NS = Op::NSL;
  static_assert(NS == 3, "three plane slots: one barrier per layer");
  __shared__ T sN[NS][NV][SPLANE];
  __shared__ TE sE[NS][NE][TPB];
  __shared__ T sX[2][2 * NC][TPB];          // a row's iy = 1 halves, for the row above
  __shared__ T sR[2][2][2 * NC][STY];       // lane 63's x = 1 corners [iy][c, iz][row]
  __shared__ T sP[NC][STX + STY + 1];       // p of the side nodes: top row, then right column (rows 0..4)
  __shared__ T sC[NC][STX + STY + 1];       // their z-carries, same order
  const int tid = threadIdx.x;
  // a wave is one row: ty is uniform across it (a scalar, so that the row
  // tests below are scalar branches)
  const int tx = tid % STX, ty = __builtin_amdgcn_readfirstlane(tid / STX);
  const int t = xcd_tile(bk.x, bk.gx);
  const int kx = t % S.ntx, ky = (t / S.ntx) % S.nty;
  const int ox = kx * STX, oy = ky * STY;
  const int z0 = (t / (S.ntx * S.nty)) * S.zc;
  const int zend = min(z0 + S.zc, L.nz);
  const int ex = ox + tx, ey = oy + ty;
  const bool exy = (ex <= L.nx - 2) && (ey <= L.ny - 2);
  const typename Op::State st = op.state();

  const bool has2 = tid + TPB < SPLANE;
  const int q1 = tid, q2 = tid + TPB;
  const int gx1 = ox + q1 % SPX, gy1 = oy + q1 / SPX;
  const int gx2 = ox + q2 % SPX, gy2 = oy + q2 / SPX;
  const bool in1 = gx1 < L.nx && gy1 < L.ny;
  const bool in2 = has2 && gx2 < L.nx && gy2 < L.ny;
  typename Op::Raw r1, r2;
  TE re[NE];
  auto fetch = [&](int pz) {
    const bool zok = pz >= 0 && pz < L.nz && pz <= zend;
    op.efetch(L, ex, ey, pz, exy && zok && pz < zend && pz <= L.nz - 2, re);
    op.fetch(L, gx1, gy1, pz, in1 && zok, st, r1);
    op.fetch(L, gx2, gy2, pz, in2 && zok, st, r2);
  };
  auto commit = [&](int slot) {
#pragma unroll
    for (int k = 0; k < NE; k++) sE[slot][k][tid] = re[k];
    T v[NV];
    op.commit(r1, st, v);
#pragma unroll
    for (int k = 0; k < NV; k++) sN[slot][k][q1] = v[k];
    if (has2) {
      op.commit(r2, st, v);
#pragma unroll
      for (int k = 0; k < NV; k++) sN[slot][k][q2] = v[k];
    }
  };
  auto slot = [](int p) { return (p + 3) % 3; };

  // Side nodes: rows 1..3's waves each take one component c = ty - 1 of the
  // tile's top row (tx, 4), on every lane; row 3's wave also takes the right
  // column (64, r) and the corner (64, 4), all three components in one pass
  // over lanes 8 c + r (r = 0..4).  Row 0's wave commits the plane's nodes
  // past the first 256 (row 1's its last 5) and takes no side work.  So every
  // wave's work per layer is about one full-width pass beside its element.
  // Lane 63 of each row hands its x = 1 corners over through LDS.  The
  // offsets (side arrays, with the component folded in, in elements; node
  // fields in bytes) are set here at plane 0 with their steps per plane; a
  // store at plane kz adds kz steps.
  const bool right = (tx == STX - 1);
  const int X = ox + tx, Y = oy + ty;
  const bool own_in = X < L.nx && Y < L.ny;
  const int sc = ty - 1;  // this wave's top-row component (row 0: none)
  const bool side_wave = ty >= 1;
  // right column and corner: row 3's wave, all three components in one
  // instruction stream, component rcc on lanes 8 rcc .. 8 rcc + 4
  const int rr = tx & 7, rcc = tx >> 3;
  const bool rc_lane = ty == STY - 1 && rr <= STY && rcc < NC;
  const bool t_in = side_wave && X < L.nx && oy + STY < L.ny;
  const bool t_own_p = oy + STY == L.ny - 1;
  const unsigned t_base = 3u * S.nr + (unsigned)sc * S.nt + (unsigned)ky * L.nx + X;
  const unsigned t_step = (unsigned)S.nty * L.nx;
  const unsigned t_ob = nid32(L, X, oy + STY, 0) * 8u;
  const int rx = ox + STX, ry = oy + rr;  // rc lanes: (64, rr), rr = 4 the corner
  const bool r_in = rc_lane && rx < L.nx && ry < L.ny;
  const bool r_own_p = rr < STY ? rx == L.nx - 1 : (rx == L.nx - 1 && ry == L.ny - 1);
  const unsigned r_base = rr < STY ? (unsigned)rcc * S.nr + (unsigned)ry * S.ntx + kx
                                   : 3u * (S.nr + S.nt) + (unsigned)rcc * S.nc + (unsigned)ky * S.ntx + kx;
  const unsigned r_step = rr < STY ? (unsigned)L.ny * S.ntx : (unsigned)S.nty * S.ntx;
  const unsigned r_ob = nid32(L, rx, ry, 0) * 8u;
  const unsigned plane8 = (unsigned)L.nx * L.ny * 8u;

  typename Op::Acc acc;
  typename CarryOf<Op>::type up{};
  T carry[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) carry[c] = T(0);

  fetch(z0 - 1);
  commit(slot(z0 - 1));
  fetch(z0);
  commit(slot(z0));
  fetch(z0 + 1);
  commit(slot(z0 + 1));
  fetch(z0 + 2);
  __syncthreads();
  auto layer = [&](int kz, auto first) {
    const int s0 = slot(kz);
    const int xb = kz & 1;
    T f[NC][2][2][2];
    {
      TE e[NE];
#pragma unroll
      for (int k = 0; k < NE; k++) e[k] = sE[s0][k][tid];
      op.element_planes(L, sN, s0, tx, ty, ex, ey, kz, st, e, f, up, decltype(first)::value);
    }
    T g[NC][2][2];
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
      for (int iz = 0; iz < 2; iz++)
#pragma unroll
        for (int iy = 0; iy < 2; iy++) g[c][iz][iy] = f[c][iz][iy][0] + wave_shr1(f[c][iz][iy][1]);
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
      for (int iz = 0; iz < 2; iz++) sX[xb][c * 2 + iz][tid] = g[c][iz][1];
    if (right) {
#pragma unroll
      for (int c = 0; c < NC; c++)
#pragma unroll
        for (int iz = 0; iz < 2; iz++)
#pragma unroll
          for (int iy = 0; iy < 2; iy++) sR[xb][iy][c * 2 + iz][ty] = f[c][iz][iy][1];
    }
    // p of this thread's nodes of plane kz, read before the slot is reused
    T own[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) own[c] = sN[s0][c][tx + SPX * ty];
    if (side_wave) sP[sc][tx] = sN[s0][sc][tx + SPX * STY];
    if (rc_lane) sP[rcc][STX + rr] = sN[s0][rcc][STX + SPX * rr];
    __syncthreads();
    commit(s0);
    fetch(kz + NS + 1);
    // node (tx, ty): the row below's iy = 1 halves come through LDS (row 0
    // has none in this tile)
    {
      T vb[NC];
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const T b0 = ty > 0 ? sX[xb][c * 2 + 0][tid - STX] : T(0);
        const T b1 = ty > 0 ? sX[xb][c * 2 + 1][tid - STX] : T(0);
        vb[c] = carry[c] + g[c][0][0] + b0;
        carry[c] = g[c][1][0] + b1;
      }
      if (own_in && kz >= z0) op.epilogue_split(L, own, X, Y, kz, vb, acc);
    }
    // side nodes, component sc: the top row (tx, 4) from row 3's iy = 1
    // halves, x-combined; the right column (64, r) from lane 63's x = 1
    // corners of row r (iy = 0) and row r - 1 (iy = 1), the corner (64, 4)
    // from row 3's
    if (side_wave) {
      const int c = sc;
      const T vt = sC[c][tx] + sX[xb][c * 2 + 0][(STY - 1) * STX + tx];
      sC[c][tx] = sX[xb][c * 2 + 1][(STY - 1) * STX + tx];
      if (t_in && kz >= z0)
        op.side1(S, (t_base + (unsigned)kz * t_step) * 8u, sP[c][tx], vt, t_own_p, t_ob + (unsigned)kz * plane8, c,
                 acc);
    }
    if (rc_lane) {
      const int c = rcc, r = rr;
      const T a0 = r < STY ? sR[xb][0][c * 2 + 0][r] : T(0), a1 = r < STY ? sR[xb][0][c * 2 + 1][r] : T(0);
      const T b0 = r > 0 ? sR[xb][1][c * 2 + 0][r - 1] : T(0), b1 = r > 0 ? sR[xb][1][c * 2 + 1][r - 1] : T(0);
      const T vr = sC[c][STX + r] + a0 + b0;
      sC[c][STX + r] = a1 + b1;
      if (r_in && kz >= z0)
        op.side1(S, (r_base + (unsigned)kz * r_step) * 8u, sP[c][STX + r], vr, r_own_p,
                 r_ob + (unsigned)kz * plane8, c, acc);
    }
  };
  layer(z0 - 1, std::true_type());
  for (int kz = z0; kz < zend; ++kz) layer(kz, std::false_type());
  op.finish(acc, bk);
}

template <class Op>
__global__ HPCR_TILE_ATTR __launch_bounds__(TPB, Op::MIN_WAVES) void tile_split_kernel(const LevelGeom L, const Op op,
                                                                                      const SplitW S) {
  tile_body_split<Op>(L, op, S, Blk::hw());
}

}  // namespace hpcr
