// Mixed-precision form of the mechanical solve: an FP64 flexible conjugate
// gradient (FCG(1), one-term orthogonalisation) on the finest level whose
// preconditioner is one FP32 multigrid V-cycle.
//
// Preconditioner.  Correction scheme on the same level hierarchy and
// rediscretised operators as the FAS solve (finest: nodal K, G, d with the
// split flag; coarse: element K_e, G_e), all vectors and coefficients in
// FP32, Chebyshev-Jacobi smoothing of a fixed degree before and after the
// coarse correction (symmetric, so the V-cycle is a fixed SPD operator up to
// the inexact coarsest solve), the coarsest level solved by FP64 Jacobi-PCG
// in one workgroup.  Chebyshev bounds come from a power iteration on D^-1 A
// per level, once per solve.
//
// Every smoother step is one tile pass: the stencil of the previous
// direction s gives A s, and the epilogue updates r -= A s, the new direction
// d = c_d s + c_r D^-1 r and the iterate x += d.  Restriction starts the
// coarse level's smoother (d_0 = D^-1 r / theta, x = d_0); prolongation
// writes the coarse correction c = P e into the direction buffer, so that the
// first post-smoothing step is the same pass with c_d = 0.
//
// Outer iteration (FP64): beta = -(z.q_prev)/(p_prev.q_prev) with z the
// preconditioner output, p = z + beta p_prev, q = A p, alpha = (p.r)/(p.q),
// u += alpha p, r -= alpha q; the residual is recomputed from u every few
// iterations and before convergence is accepted.
#include <hip/hip_runtime.h>
#include <type_traits>

#include "hpcr_internal.h"
#include "hpcr_tile.h"
#include "hpcr_multi.h"

namespace hpcr {

// F_CHEBQ: a Chebyshev step that also accumulates x.q_prev into S_ZQ
enum FMode { F_CHEB = 0, F_RES = 1, F_POWER = 2, F_CHEBQ = 3 };

template <int MODE, bool FINE>
struct MechOpF {
  using T = float;
  using TR = float;
  static constexpr int NC = 3;
  static constexpr int NU = 3;
  static constexpr int NV = NU;
  static constexpr int NR = NV;
  static constexpr bool NEEDS_OWN = (MODE != F_RES);
  static constexpr int NSL = HPCR_MIXF_SLOTS, MIN_WAVES = HPCR_MIXF_WAVES;
  using Raw = float[NR];
  struct Acc {
    double a[2];
    __device__ Acc() { a[0] = a[1] = 0.0; }
  };
  struct State {};
  // finest level: (lambda, mu) at the 8 Gauss points, precomputed once per
  // solve; coarse levels: K_e, G_e
  // finest level: four float4 per element, (lambda, mu) of Gauss points
  // 2k and 2k+1 in float4 k (16-byte loads and LDS accesses)
  using TE = std::conditional_t<FINE, float4, float>;
  static constexpr int NE = FINE ? 4 : 2;
  MixCoef C;
  const float* s;        // stencil vector
  float* dout;           // CHEB: new direction; POWER: D^-1 A s
  float* r;              // CHEB, RES: residual, updated in place
  float* x;              // CHEB: iterate, updated in place
  const float* invD;
  float cd, cr;          // CHEB: d = cd s + cr D^-1 r
  const double* qprev;   // CHEB (optional): accumulate x.qprev into S_ZQ
  float sx, sy, sz, tx_, ty_, tz_;
  Reduce R;
  size_t N;

  __device__ __forceinline__ State state() const { return State{}; }

  __device__ __forceinline__ void fetch(const LevelGeom& L, int gx, int gy, int gz, bool ok, const State&,
                                        float (&v)[NR]) const {
#pragma unroll
    for (int k = 0; k < NR; k++) v[k] = 0.0f;
    if (!ok) return;
    const size_t q = nidx(L, gx, gy, gz);
#pragma unroll
    for (int c = 0; c < 3; c++) v[c] = s[c * N + q];
  }
  __device__ __forceinline__ void efetch(const LevelGeom& L, int ex, int ey, int ez, bool ok, TE (&e)[NE]) const {
#pragma unroll
    for (int k = 0; k < NE; k++) e[k] = TE{};
    if (!ok) return;
    const size_t q = nidx(L, ex, ey, ez);
    if constexpr (FINE) {
      const float4* p = reinterpret_cast<const float4*>(C.lm) + 4 * q;
#pragma unroll
      for (int k = 0; k < 4; k++) e[k] = p[k];
    } else {
      e[0] = C.Ke[q];
      e[1] = C.Ge[q];
    }
  }
  __device__ __forceinline__ void commit(const float (&v)[NR], const State&, float (&o)[NV]) const {
#pragma unroll
    for (int k = 0; k < NV; k++) o[k] = v[k];
  }

  __device__ __forceinline__ void element(const LevelGeom& L, const float (&sN)[NSL][NV][PLANE], int s0, int tx,
                                          int ty, int ex, int ey, int ez, const State&, const TE (&e)[NE],
                                          float (&f)[3][2][2][2]) const {
    float u[3][2][2][2];
#pragma unroll
    for (int c = 0; c < 3; c++) gather8(sN, s0, tx, ty, c, u[c]);
    if constexpr (FINE) {
      float lam[2][2][2], mu[2][2][2];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int g0 = 2 * k, g1 = 2 * k + 1;
        lam[g0 >> 2][(g0 >> 1) & 1][g0 & 1] = e[k].x;
        mu[g0 >> 2][(g0 >> 1) & 1][g0 & 1] = e[k].y;
        lam[g1 >> 2][(g1 >> 1) & 1][g1 & 1] = e[k].z;
        mu[g1 >> 2][(g1 >> 1) & 1][g1 & 1] = e[k].w;
      }
      mech_element<false>(u, lam, mu, 0.0f, 0.0f, sx, sy, sz, tx_, ty_, tz_, f);
    } else {
      const float lam0 = e[0] - (2.0f / 3.0f) * e[1];
      float dl[2][2][2], dm[2][2][2];
      mech_element<true>(u, dl, dm, lam0, e[1], sx, sy, sz, tx_, ty_, tz_, f);
    }
  }

  // the epilogue's inputs at the owner's node, read one plane ahead
  static constexpr bool CHEB = (MODE == F_CHEB || MODE == F_CHEBQ);
  struct NData {
    float r[3], x[3], iD[3];
    double qp[MODE == F_CHEBQ ? 3 : 1];
  };
  __device__ __forceinline__ void nfetch(const LevelGeom& L, int i, int j, int k, NData& n) const {
    const size_t q = nidx(L, i, j, k);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      if constexpr (MODE != F_POWER) n.r[c] = r[c * N + q];
      if constexpr (CHEB) n.x[c] = x[c * N + q];
      if constexpr (MODE != F_RES) n.iD[c] = invD[c * N + q];
      if constexpr (MODE == F_CHEBQ) n.qp[c] = qprev[c * N + q];
    }
  }

  __device__ __forceinline__ void epilogue(const LevelGeom& L, const float (&own)[3], int i, int j, int k, int dm,
                                           const float (&v)[3], const State&, const NData& nd,
                                           Acc& acc) const {
    const size_t q = nidx(L, i, j, k);
    if constexpr (CHEB) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const size_t o = c * N + q;
        const float rn = ((dm >> c) & 1) ? 0.0f : nd.r[c] - v[c];
        const float dn = cd * own[c] + cr * nd.iD[c] * rn;
        const float xn = nd.x[c] + dn;
        r[o] = rn;
        dout[o] = dn;
        x[o] = xn;
        if constexpr (MODE == F_CHEBQ) acc.a[0] += (double)xn * nd.qp[c];
      }
    } else if constexpr (MODE == F_RES) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const size_t o = c * N + q;
        r[o] = ((dm >> c) & 1) ? 0.0f : nd.r[c] - v[c];
      }
    } else {  // F_POWER
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const size_t o = c * N + q;
        const bool dir = (dm >> c) & 1;
        const float di = nd.iD[c];
        dout[o] = dir ? 0.0f : di * v[c];
        if (!dir) {
          acc.a[0] += (double)own[c] * (double)v[c];
          acc.a[1] += (double)own[c] * (double)own[c] / (double)di;
        }
      }
    }
  }

  __device__ __forceinline__ void finish(Acc& acc, const Blk& bk) const {
    double* scal = R.scal;
    if constexpr (MODE == F_CHEBQ) {
      double v1[1] = {acc.a[0]};
      reduce_finalize<RED_SUM, 1>(bk, v1, R, 0.0, [=](double (&t)[1]) { scal[S_ZQ] = t[0]; });
    } else if constexpr (MODE == F_POWER) {
      reduce_finalize<RED_SUM, 2>(bk, acc.a, R, 0.0, [=](double (&t)[2]) { scal[S_LAM] = t[0] / t[1]; });
    }
  }
};

// FCG pass A (FP64, finest level): p = z + beta p_prev with z the FP32
// preconditioner output; q = A p; p.q, p.r -> alpha.  LAG: also
// u += alpha_prev p_prev at the own nodes, the update of the previous
// iteration that its lean pass B left out (the same fma, one kernel later).
template <bool LAG>
struct FcgOpA {
  using T = double;
  using TR = double;
  static constexpr int NC = 3;
  static constexpr int NU = 3;
  static constexpr int NV = 6;
  static constexpr int NR = 9;
  static constexpr bool NEEDS_OWN = true;
  static constexpr int NSL = NSLOT, MIN_WAVES = 1;
  // raw node values in flight: z stays FP32 until the commit, so that no
  // conversion consumes a load before the layer that needs it
  struct Raw {
    float z[3];
    double p[3], K, G, d;
  };
  struct Acc {
    double a[2];
    __device__ Acc() { a[0] = a[1] = 0.0; }
  };
  struct State {
    double beta, aprev;
  };
  // the element's split flag, staged through LDS with the node planes
  using TE = int;
  static constexpr int NE = 1;
  MechCoef C;
  const float* z;
  const double* p_old;
  double* p_new;
  double* q;
  const double* r;
  double* u;  // LAG
  int first;
  double sx, sy, sz, tx_, ty_, tz_;
  Reduce R;
  size_t N;

  __device__ __forceinline__ State state() const {
    State st;
    st.beta = first ? 0.0 : -R.scal[S_ZQ] / R.scal[S_PQ];
    st.aprev = LAG ? R.scal[S_APREV] : 0.0;
    return st;
  }
  __device__ __forceinline__ void fetch(const LevelGeom& L, int gx, int gy, int gz, bool ok, const State&,
                                        Raw& v) const {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      v.z[c] = 0.0f;
      v.p[c] = 0.0;
    }
    v.K = v.G = v.d = 0.0;
    if (!ok) return;
    const size_t o = nidx(L, gx, gy, gz);
#pragma unroll
    for (int c = 0; c < 3; c++) v.z[c] = z[c * N + o];
    if (!first) {
#pragma unroll
      for (int c = 0; c < 3; c++) v.p[c] = p_old[c * N + o];
    }
    v.K = C.K[o];
    v.G = C.G[o];
    v.d = C.d[o];
  }
  __device__ __forceinline__ void efetch(const LevelGeom& L, int ex, int ey, int ez, bool ok, int (&e)[1]) const {
    e[0] = ok ? (int)C.flag[nidx(L, ex, ey, ez)] : 0;
  }
  __device__ __forceinline__ void commit(const Raw& v, const State& st, double (&w)[NV]) const {
#pragma unroll
    for (int c = 0; c < 3; c++) w[c] = (double)v.z[c] + st.beta * v.p[c];
    w[3] = v.K;
    w[4] = v.G;
    w[5] = v.d;
  }
  __device__ __forceinline__ void element(const LevelGeom& L, const double (&sN)[NSL][NV][PLANE], int s0, int tx,
                                          int ty, int ex, int ey, int ez, const State&, const int (&e)[1],
                                          double (&f)[3][2][2][2]) const {
    double u[3][2][2][2], Kn[2][2][2], Gn[2][2][2], dn[2][2][2], lam[2][2][2], mu[2][2][2];
#pragma unroll
    for (int c = 0; c < 3; c++) gather8(sN, s0, tx, ty, c, u[c]);
    gather8(sN, s0, tx, ty, 3, Kn);
    gather8(sN, s0, tx, ty, 4, Gn);
    gather8(sN, s0, tx, ty, 5, dn);
    fine_lame(Kn, Gn, dn, e[0], C.small_k, lam, mu);
    mech_element<false>(u, lam, mu, 0.0, 0.0, sx, sy, sz, tx_, ty_, tz_, f);
  }
  struct NData {
    double r[3];
    double u[LAG ? 3 : 1], p[LAG ? 3 : 1];
  };
  __device__ __forceinline__ void nfetch(const LevelGeom& L, int i, int j, int k, NData& n) const {
    const size_t o = nidx(L, i, j, k);
#pragma unroll
    for (int c = 0; c < 3; c++) n.r[c] = r[c * N + o];
    if constexpr (LAG) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        n.u[c] = u[c * N + o];
        n.p[c] = p_old[c * N + o];
      }
    }
  }
  __device__ __forceinline__ void epilogue(const LevelGeom& L, const double (&own)[3], int i, int j, int k, int dm,
                                           const double (&v)[3], const State& st, const NData& nd,
                                           Acc& acc) const {
    const size_t o = nidx(L, i, j, k);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      p_new[c * N + o] = own[c];
      q[c * N + o] = v[c];
      acc.a[0] += own[c] * v[c];
      acc.a[1] += own[c] * nd.r[c];
      if constexpr (LAG) u[c * N + o] = fma(st.aprev, nd.p[c], nd.u[c]);
    }
  }
  __device__ __forceinline__ void finish(Acc& acc, const Blk& bk) const {
    double* scal = R.scal;
    reduce_finalize<RED_SUM, 2>(bk, acc.a, R, 0.0, [=](double (&t)[2]) {
      scal[S_PQ] = t[0];
      scal[S_PR] = t[1];
      scal[S_FALPHA] = t[1] / t[0];
    });
  }
};

static dim3 mix_tile_grid(const LevelGeom& L) { return dim3((unsigned)tile_blocks(L.nx, L.ny, L.nz, L.zc)); }

template <int MODE, bool FINE>
static void run_f32(hipStream_t st, const LevelGeom& L, const MixCoef& C, const float* s, float* dout, float* r,
                    float* x, const float* invD, float cd, float cr, const double* qprev, const Reduce& R) {
  MechOpF<MODE, FINE> op{};
  op.C = C;
  op.s = s;
  op.dout = dout;
  op.r = r;
  op.x = x;
  op.invD = invD;
  op.cd = cd;
  op.cr = cr;
  op.qprev = qprev;
  const double wg = 0.125 * L.hx * L.hy * L.hz;
  op.sx = (float)(1.0 / L.hx);
  op.sy = (float)(1.0 / L.hy);
  op.sz = (float)(1.0 / L.hz);
  op.tx_ = (float)(wg / L.hx);
  op.ty_ = (float)(wg / L.hy);
  op.tz_ = (float)(wg / L.hz);
  op.R = R;
  op.N = (size_t)L.nx * L.ny * L.nz;
  if (record<TileK<MechOpF<MODE, FINE>>>(mix_tile_grid(L), {L, op})) return;
  tile_kernel<MechOpF<MODE, FINE>><<<mix_tile_grid(L), dim3(TPB), 0, st>>>(L, op);
}

template <int MODE>
static void run_f32_level(hipStream_t st, const LevelGeom& L, const MixCoef& C, const float* s, float* dout,
                          float* r, float* x, const float* invD, float cd, float cr, const double* qprev,
                          const Reduce& R) {
  if (C.fine)
    run_f32<MODE, true>(st, L, C, s, dout, r, x, invD, cd, cr, qprev, R);
  else
    run_f32<MODE, false>(st, L, C, s, dout, r, x, invD, cd, cr, qprev, R);
}

void launch_mix_cheb(hipStream_t st, const LevelGeom& L, const MixCoef& C, const float* s, float* dout, float* r,
                     float* x, const float* invD, float cd, float cr, const double* qprev, const Reduce& R) {
  if (qprev)
    run_f32_level<F_CHEBQ>(st, L, C, s, dout, r, x, invD, cd, cr, qprev, R);
  else
    run_f32_level<F_CHEB>(st, L, C, s, dout, r, x, invD, cd, cr, nullptr, R);
}
void launch_mix_res(hipStream_t st, const LevelGeom& L, const MixCoef& C, const float* s, float* r) {
  Reduce R{};
  run_f32_level<F_RES>(st, L, C, s, nullptr, r, nullptr, nullptr, 0.0f, 0.0f, nullptr, R);
}
void launch_mix_power(hipStream_t st, const LevelGeom& L, const MixCoef& C, const float* v, float* w,
                      const float* invD, const Reduce& R) {
  run_f32_level<F_POWER>(st, L, C, v, w, nullptr, nullptr, invD, 0.0f, 0.0f, nullptr, R);
}

template <bool LAG>
static void run_fcg_a(hipStream_t st, const LevelGeom& L, const MechCoef& C, const float* z, const double* p_old,
                      double* p_new, double* q, const double* r, int first, double* u, const Reduce& R) {
  FcgOpA<LAG> op{};
  op.C = C;
  op.u = u;
  op.z = z;
  op.p_old = p_old;
  op.p_new = p_new;
  op.q = q;
  op.r = r;
  op.first = first;
  const double wg = 0.125 * L.hx * L.hy * L.hz;
  op.sx = 1.0 / L.hx;
  op.sy = 1.0 / L.hy;
  op.sz = 1.0 / L.hz;
  op.tx_ = wg / L.hx;
  op.ty_ = wg / L.hy;
  op.tz_ = wg / L.hz;
  op.R = R;
  op.N = (size_t)L.nx * L.ny * L.nz;
  if (record<TileK<FcgOpA<LAG>>>(mix_tile_grid(L), {L, op})) return;
  tile_kernel<FcgOpA<LAG>><<<mix_tile_grid(L), dim3(TPB), 0, st>>>(L, op);
}
void launch_fcg_a(hipStream_t st, const LevelGeom& L, const MechCoef& C, const float* z, const double* p_old,
                  double* p_new, double* q, const double* r, int first, double* u_lag, const Reduce& R) {
  if (u_lag && !first)
    run_fcg_a<true>(st, L, C, z, p_old, p_new, q, r, 0, u_lag, R);
  else
    run_fcg_a<false>(st, L, C, z, p_old, p_new, q, r, first, nullptr, R);
}

// ---------------------------------------------------------------------------
// pointwise kernels
// ---------------------------------------------------------------------------
constexpr int MX_TPB = 256;
static int mx_blocks(size_t n, size_t cap) {
  const size_t b = (n + MX_TPB - 1) / MX_TPB;
  return (int)(b == 0 ? 1 : (b < cap ? b : cap));
}

__device__ __forceinline__ int node_mask(const LevelGeom& L, unsigned q) {
  const unsigned plane = (unsigned)L.nx * (unsigned)L.ny;
  const unsigned k = q / plane;
  if (!L.fz[k]) return 0;
  const unsigned rem = q - k * plane;
  const unsigned j = rem / (unsigned)L.nx, i = rem - j * (unsigned)L.nx;
  return dir_mask(L, (int)i, (int)j, (int)k);
}

// start of the finest-level smoother from r (FP64): r32 = r, d = cr0 D^-1 r, x = d
__global__ __launch_bounds__(MX_TPB) void mix_start_kernel(unsigned N, const double* __restrict__ r64,
                                                           float* __restrict__ r32, float* __restrict__ d,
                                                           float* __restrict__ x, const float* __restrict__ invD,
                                                           float cr0) {
  for (unsigned q = blockIdx.x * MX_TPB + threadIdx.x; q < N; q += gridDim.x * MX_TPB) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const size_t o = (size_t)c * N + q;
      const float rv = (float)r64[o];
      const float dv = cr0 * invD[o] * rv;
      r32[o] = rv;
      d[o] = dv;
      x[o] = dv;
    }
  }
}

// FCG pass B (FP64): u += alpha p; r -= alpha q (masked); r.r; then the start
// of the next preconditioner application from the new r.  LDSF: the
// Dirichlet axis flags are staged in LDS (nx + ny + nz <= DIRF_MAX).
// LEAN: u and p are left alone (the next pass A, or launch_fcg_flush,
// applies u += alpha p) and alpha moves to scal[S_APREV].
template <bool LDSF, bool LEAN>
__device__ __forceinline__ void fcg_b_body(const LevelGeom& L, double* __restrict__ u, double* __restrict__ r,
                                           const double* __restrict__ p, const double* __restrict__ q,
                                           float* __restrict__ r32, float* __restrict__ d, float* __restrict__ x,
                                           const float* __restrict__ invD, float cr0, const Reduce& R,
                                           const Blk& bk) {
  __shared__ unsigned char sflags[LDSF ? DIRF_MAX : 1];
  const DirFlags F{sflags};
  if constexpr (LDSF) {
    F.stage(L);
    __syncthreads();
  }
  const unsigned N = (unsigned)Nof(L);
  const unsigned plane = (unsigned)L.nx * (unsigned)L.ny;
  const double alpha = R.scal[S_FALPHA];
  double acc[1] = {0.0};
  for (unsigned n = bk.x * MX_TPB + threadIdx.x; n < N; n += bk.gx * MX_TPB) {
    int dm;
    if constexpr (LDSF) {
      const unsigned k = n / plane, rem = n - k * plane;
      const unsigned j = rem / (unsigned)L.nx, i = rem - j * (unsigned)L.nx;
      dm = F.mask(L, (int)i, (int)j, (int)k);
    } else {
      dm = node_mask(L, n);
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const size_t o = (size_t)c * N + n;
      if constexpr (!LEAN) u[o] = fma(alpha, p[o], u[o]);
      double rv = r[o] - alpha * q[o];
      if ((dm >> c) & 1) rv = 0.0;
      r[o] = rv;
      acc[0] += rv * rv;
      const float rf = (float)rv;
      const float dv = cr0 * invD[o] * rf;
      r32[o] = rf;
      d[o] = dv;
      x[o] = dv;
    }
  }
  double* scal = R.scal;
  reduce_finalize<RED_SUM, 1>(bk, acc, R, 0.0, [=](double (&t)[1]) {
    scal[S_RR] = t[0];
    if constexpr (LEAN) scal[S_APREV] = scal[S_FALPHA];
  });
}
template <bool LDSF, bool LEAN>
__global__ __launch_bounds__(MX_TPB) void fcg_b_kernel(const LevelGeom L, double* __restrict__ u,
                                                       double* __restrict__ r, const double* __restrict__ p,
                                                       const double* __restrict__ q, float* __restrict__ r32,
                                                       float* __restrict__ d, float* __restrict__ x,
                                                       const float* __restrict__ invD, float cr0, const Reduce R) {
  fcg_b_body<LDSF, LEAN>(L, u, r, p, q, r32, d, x, invD, cr0, R, Blk::hw());
}
struct FcgBArgs {
  LevelGeom L;
  double* u;
  double* r;
  const double* p;
  const double* q;
  float* r32;
  float* d;
  float* x;
  const float* invD;
  float cr0;
  Reduce R;
};
template <bool LDSF, bool LEAN>
struct FcgBK {
  using Args = FcgBArgs;
  static constexpr int NT = MX_TPB, MINW = 1;
  static dim3 block() { return dim3(MX_TPB); }
  __device__ __forceinline__ static void run(const Args& a, const Blk& bk) {
    fcg_b_body<LDSF, LEAN>(a.L, a.u, a.r, a.p, a.q, a.r32, a.d, a.x, a.invD, a.cr0, a.R, bk);
  }
};

__global__ __launch_bounds__(MX_TPB) void fcg_flush_kernel(size_t n, double* __restrict__ u,
                                                           const double* __restrict__ p,
                                                           const double* __restrict__ scal) {
  const double alpha = scal[S_APREV];
  for (size_t o = (size_t)blockIdx.x * MX_TPB + threadIdx.x; o < n; o += (size_t)gridDim.x * MX_TPB)
    u[o] = fma(alpha, p[o], u[o]);
}

// restriction of the FP32 residual (full weighting, as the FP64 transfer);
// starts the coarse smoother (d = cr0 D^-1 r, x = d), or, on the coarsest
// level (start == 0), only clears x for the coarsest solve
__device__ __forceinline__ void mix_ijk(const LevelGeom& L, unsigned q, int& i, int& j, int& k) {
  const unsigned plane = (unsigned)L.nx * (unsigned)L.ny;
  const unsigned kk = q / plane, rem = q - kk * plane;
  const unsigned jj = rem / (unsigned)L.nx;
  i = (int)(rem - jj * (unsigned)L.nx);
  j = (int)jj;
  k = (int)kk;
}

__device__ __forceinline__ void mix_restrict_body(const LevelGeom& Lc, const LevelGeom& Lf,
                                                  const float* __restrict__ rf, float* __restrict__ rc,
                                                  float* __restrict__ dc, float* __restrict__ xc,
                                                  const float* __restrict__ invD, float cr0, int start, const Blk& bk) {
  const unsigned qn = bk.x * 256 + threadIdx.x;
  if (qn >= (unsigned)Nof(Lc)) return;
  int i, j, k;
  mix_ijk(Lc, qn, i, j, k);
  const size_t Nc = Nof(Lc), Nf = Nof(Lf);
  const int x = 2 * i, y = 2 * j, z = 2 * k;
  float wx[3], wy[3], wz[3];
#pragma unroll
  for (int t = 0; t < 3; t++) {
    wx[t] = (x + t - 1 < 0 || x + t - 1 > Lf.nx - 1) ? 0.0f : (t == 1 ? 1.0f : 0.5f);
    wy[t] = (y + t - 1 < 0 || y + t - 1 > Lf.ny - 1) ? 0.0f : (t == 1 ? 1.0f : 0.5f);
    wz[t] = (z + t - 1 < 0 || z + t - 1 > Lf.nz - 1) ? 0.0f : (t == 1 ? 1.0f : 0.5f);
  }
  const size_t qc = nidx(Lc, i, j, k);
  const int dm = dir_mask(Lc, i, j, k);
  for (int c = 0; c < 3; c++) {
    const float* r = rf + c * Nf;
    float acc = 0.0f;
#pragma unroll
    for (int dz = 0; dz < 3; dz++) {
      if (wz[dz] == 0.0f) continue;
      float accy = 0.0f;
#pragma unroll
      for (int dy = 0; dy < 3; dy++) {
        if (wy[dy] == 0.0f) continue;
        const size_t row = nidx(Lf, 0, y + dy - 1, z + dz - 1);
        float accx = 0.0f;
#pragma unroll
        for (int dx = 0; dx < 3; dx++)
          if (wx[dx] != 0.0f) accx += wx[dx] * r[row + x + dx - 1];
        accy += wy[dy] * accx;
      }
      acc += wz[dz] * accy;
    }
    if ((dm >> c) & 1) acc = 0.0f;
    const size_t o = c * Nc + qc;
    rc[o] = acc;
    if (start) {
      const float dv = cr0 * invD[o] * acc;
      dc[o] = dv;
      xc[o] = dv;
    } else {
      xc[o] = 0.0f;
    }
  }
}
__global__ __launch_bounds__(256) void mix_restrict_kernel(const LevelGeom Lc, const LevelGeom Lf,
                                                           const float* __restrict__ rf, float* __restrict__ rc,
                                                           float* __restrict__ dc, float* __restrict__ xc,
                                                           const float* __restrict__ invD, float cr0, int start) {
  mix_restrict_body(Lc, Lf, rf, rc, dc, xc, invD, cr0, start, Blk::hw());
}
struct MixRestrictK {
  struct Args {
    LevelGeom Lc;
    LevelGeom Lf;
    const float* rf;
    float* rc;
    float* dc;
    float* xc;
    const float* invD;
    float cr0;
    int start;
  };
  static constexpr int NT = 256, MINW = 1;
  static dim3 block() { return dim3(256); }
  __device__ __forceinline__ static void run(const Args& a, const Blk& bk) {
    mix_restrict_body(a.Lc, a.Lf, a.rf, a.rc, a.dc, a.xc, a.invD, a.cr0, a.start, bk);
  }
};

// prolongation of the coarse correction: c = P e (zero on Dirichlet DOFs),
// x_f += c, c kept as the stencil vector of the first post-smoothing step
__device__ __forceinline__ void mix_prolong_body(const LevelGeom& Lf, const LevelGeom& Lc, float* __restrict__ xf,
                                                 float* __restrict__ cf, const float* __restrict__ e, const Blk& bk) {
  const unsigned qn = bk.x * 256 + threadIdx.x;
  if (qn >= (unsigned)Nof(Lf)) return;
  int x, y, z;
  mix_ijk(Lf, qn, x, y, z);
  const size_t Nc = Nof(Lc), Nf = Nof(Lf);
  const int i0 = x >> 1, j0 = y >> 1, k0 = z >> 1;
  const int i1 = i0 + (x & 1), j1 = j0 + (y & 1), k1 = k0 + (z & 1);
  const float w = ((x & 1) ? 0.5f : 1.0f) * ((y & 1) ? 0.5f : 1.0f) * ((z & 1) ? 0.5f : 1.0f);
  const size_t qf = nidx(Lf, x, y, z);
  const int dm = dir_mask(Lf, x, y, z);
  const size_t q00 = nidx(Lc, 0, j0, k0), q10 = nidx(Lc, 0, j1, k0), q01 = nidx(Lc, 0, j0, k1),
               q11 = nidx(Lc, 0, j1, k1);
  for (int c = 0; c < 3; c++) {
    const float* ec = e + c * Nc;
    const float a00 = ec[q00 + i0] + ((x & 1) ? ec[q00 + i1] : 0.0f);
    const float a10 = (y & 1) ? ec[q10 + i0] + ((x & 1) ? ec[q10 + i1] : 0.0f) : 0.0f;
    const float a01 = (z & 1) ? ec[q01 + i0] + ((x & 1) ? ec[q01 + i1] : 0.0f) : 0.0f;
    const float a11 = ((y & 1) && (z & 1)) ? ec[q11 + i0] + ((x & 1) ? ec[q11 + i1] : 0.0f) : 0.0f;
    float corr = w * (a00 + a10 + a01 + a11);
    if ((dm >> c) & 1) corr = 0.0f;
    const size_t o = c * Nf + qf;
    xf[o] += corr;
    cf[o] = corr;
  }
}
__global__ __launch_bounds__(256) void mix_prolong_kernel(const LevelGeom Lf, const LevelGeom Lc,
                                                          float* __restrict__ xf, float* __restrict__ cf,
                                                          const float* __restrict__ e) {
  mix_prolong_body(Lf, Lc, xf, cf, e, Blk::hw());
}
struct MixProlongK {
  struct Args {
    LevelGeom Lf;
    LevelGeom Lc;
    float* xf;
    float* cf;
    const float* e;
  };
  static constexpr int NT = 256, MINW = 1;
  static dim3 block() { return dim3(256); }
  __device__ __forceinline__ static void run(const Args& a, const Blk& bk) {
    mix_prolong_body(a.Lf, a.Lc, a.xf, a.cf, a.e, bk);
  }
};

// (lambda, mu) at the Gauss points of every finest-level element (the
// coefficients of MP_all_GAUSS_Point, computed in FP64 and stored as FP32)
__global__ __launch_bounds__(256) void mix_lame_kernel(const LevelGeom L, const double* __restrict__ K,
                                                       const double* __restrict__ G, const double* __restrict__ d,
                                                       const signed char* __restrict__ flag, double small_k,
                                                       float* __restrict__ lm) {
  const int ex = blockIdx.x * 64 + threadIdx.x, ey = blockIdx.y * 4 + threadIdx.y, ez = blockIdx.z;
  if (ex > L.nx - 2 || ey > L.ny - 2) return;
  double Kn[2][2][2], Gn[2][2][2], dn[2][2][2], lam[2][2][2], mu[2][2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int c = 0; c < 2; c++) {
        const size_t q = nidx(L, ex + c, ey + b, ez + a);
        Kn[a][b][c] = K[q];
        Gn[a][b][c] = G[q];
        dn[a][b][c] = d[q];
      }
  const size_t e = nidx(L, ex, ey, ez);
  fine_lame(Kn, Gn, dn, (int)flag[e], small_k, lam, mu);
  // element-major: 16 floats per element, Gauss points 2k and 2k+1 in float4 k
  float4* o = reinterpret_cast<float4*>(lm) + 4 * e;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int g0 = 2 * k, g1 = 2 * k + 1;
    o[k] = make_float4((float)lam[g0 >> 2][(g0 >> 1) & 1][g0 & 1], (float)mu[g0 >> 2][(g0 >> 1) & 1][g0 & 1],
                       (float)lam[g1 >> 2][(g1 >> 1) & 1][g1 & 1], (float)mu[g1 >> 2][(g1 >> 1) & 1][g1 & 1]);
  }
}
void launch_mix_lame(hipStream_t st, const LevelGeom& L, const double* K, const double* G, const double* d,
                     const signed char* flag, double small_k, float* lm) {
  const dim3 g((L.nx - 1 + 63) / 64, (L.ny - 1 + 3) / 4, L.nz - 1);
  hipLaunchKernelGGL(mix_lame_kernel, g, dim3(64, 4), 0, st, L, K, G, d, flag, small_k, lm);
}

__device__ __forceinline__ void to_f64_body(size_t n, const float* __restrict__ a, double* __restrict__ b,
                                            const Blk& bk) {
  for (size_t q = (size_t)bk.x * blockDim.x + threadIdx.x; q < n; q += (size_t)bk.gx * blockDim.x)
    b[q] = (double)a[q];
}
__global__ void to_f64_kernel(size_t n, const float* __restrict__ a, double* __restrict__ b) {
  to_f64_body(n, a, b, Blk::hw());
}
struct ToF64K {
  struct Args {
    size_t n;
    const float* a;
    double* b;
  };
  static constexpr int NT = 256, MINW = 1;
  static dim3 block() { return dim3(256); }
  __device__ __forceinline__ static void run(const Args& a, const Blk& bk) {
    to_f64_body(a.n, a.a, a.b, bk);
  }
};
__device__ __forceinline__ void to_f32_body(size_t n, const double* __restrict__ a, float* __restrict__ b,
                                            const Blk& bk) {
  for (size_t q = (size_t)bk.x * blockDim.x + threadIdx.x; q < n; q += (size_t)bk.gx * blockDim.x)
    b[q] = (float)a[q];
}
__global__ void to_f32_kernel(size_t n, const double* __restrict__ a, float* __restrict__ b) {
  to_f32_body(n, a, b, Blk::hw());
}
struct ToF32K {
  struct Args {
    size_t n;
    const double* a;
    float* b;
  };
  static constexpr int NT = 256, MINW = 1;
  static dim3 block() { return dim3(256); }
  __device__ __forceinline__ static void run(const Args& a, const Blk& bk) {
    to_f32_body(a.n, a.a, a.b, bk);
  }
};

// deterministic start vector of the power iteration, zero on Dirichlet DOFs
__global__ void power_init_kernel(const LevelGeom L, float* __restrict__ v) {
  const unsigned N = (unsigned)Nof(L);
  for (unsigned n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
    const int dm = node_mask(L, n);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      unsigned h = (n * 3u + (unsigned)c) * 2654435761u;
      h ^= h >> 15;
      h *= 2246822519u;
      h ^= h >> 13;
      const float val = 0.5f + (float)(h & 0xffffu) * (1.0f / 65536.0f);
      v[(size_t)c * N + n] = ((dm >> c) & 1) ? 0.0f : val;
    }
  }
}

void launch_mix_start(hipStream_t st, const LevelGeom& L, const double* r64, float* r32, float* d, float* x,
                      const float* invD, float cr0) {
  const size_t N = Nof(L);
  hipLaunchKernelGGL(mix_start_kernel, dim3(mx_blocks(N, 4096)), dim3(MX_TPB), 0, st, (unsigned)N, r64, r32, d, x,
                     invD, cr0);
}
void launch_fcg_b(hipStream_t st, const LevelGeom& L, double* u, double* r, const double* p, const double* q,
                  float* r32, float* d, float* x, const float* invD, float cr0, bool lean, const Reduce& R) {
  const dim3 g(passb_blocks(Nof(L))), b(MX_TPB);
  const bool ldsf = L.nx + L.ny + L.nz <= DIRF_MAX;
  const FcgBArgs a{L, u, r, p, q, r32, d, x, invD, cr0, R};
  bool rec;
  if (ldsf && lean)
    rec = record<FcgBK<true, true>>(g, a);
  else if (ldsf)
    rec = record<FcgBK<true, false>>(g, a);
  else if (lean)
    rec = record<FcgBK<false, true>>(g, a);
  else
    rec = record<FcgBK<false, false>>(g, a);
  if (rec) return;
  if (ldsf && lean)
    hipLaunchKernelGGL((fcg_b_kernel<true, true>), g, b, 0, st, L, u, r, p, q, r32, d, x, invD, cr0, R);
  else if (ldsf)
    hipLaunchKernelGGL((fcg_b_kernel<true, false>), g, b, 0, st, L, u, r, p, q, r32, d, x, invD, cr0, R);
  else if (lean)
    hipLaunchKernelGGL((fcg_b_kernel<false, true>), g, b, 0, st, L, u, r, p, q, r32, d, x, invD, cr0, R);
  else
    hipLaunchKernelGGL((fcg_b_kernel<false, false>), g, b, 0, st, L, u, r, p, q, r32, d, x, invD, cr0, R);
}
void launch_fcg_flush(hipStream_t st, const LevelGeom& L, double* u, const double* p, const Reduce& R) {
  const size_t n = 3 * Nof(L);
  hipLaunchKernelGGL(fcg_flush_kernel, dim3(mx_blocks(n, 4096)), dim3(MX_TPB), 0, st, n, u, p,
                     static_cast<const double*>(R.scal));
}
void launch_mix_restrict(hipStream_t st, const LevelGeom& Lc, const LevelGeom& Lf, const float* rf, float* rc,
                         float* dc, float* xc, const float* invD, float cr0, int start) {
  if (record<MixRestrictK>(dim3((unsigned)((Nof(Lc) + 255) / 256)), {Lc, Lf, rf, rc, dc, xc, invD, cr0, start}))
    return;
  hipLaunchKernelGGL(mix_restrict_kernel, dim3((unsigned)((Nof(Lc) + 255) / 256)), dim3(256), 0, st, Lc, Lf, rf, rc,
                     dc, xc, invD, cr0, start);
}
void launch_mix_prolong(hipStream_t st, const LevelGeom& Lf, const LevelGeom& Lc, float* xf, float* cf,
                        const float* ec) {
  if (record<MixProlongK>(dim3((unsigned)((Nof(Lf) + 255) / 256)), {Lf, Lc, xf, cf, ec})) return;
  hipLaunchKernelGGL(mix_prolong_kernel, dim3((unsigned)((Nof(Lf) + 255) / 256)), dim3(256), 0, st, Lf, Lc, xf, cf,
                     ec);
}
void launch_to_f32(hipStream_t st, size_t n, const double* a, float* b) {
  if (record<ToF32K>(dim3(mx_blocks(n, 4096)), {n, a, b})) return;
  hipLaunchKernelGGL(to_f32_kernel, dim3(mx_blocks(n, 4096)), dim3(256), 0, st, n, a, b);
}
void launch_to_f64(hipStream_t st, size_t n, const float* a, double* b) {
  if (record<ToF64K>(dim3(mx_blocks(n, 4096)), {n, a, b})) return;
  hipLaunchKernelGGL(to_f64_kernel, dim3(mx_blocks(n, 4096)), dim3(256), 0, st, n, a, b);
}
void launch_power_init(hipStream_t st, const LevelGeom& L, float* v) {
  hipLaunchKernelGGL(power_init_kernel, dim3(mx_blocks(Nof(L), 4096)), dim3(256), 0, st, L, v);
}

}  // namespace hpcr
