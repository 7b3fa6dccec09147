// Mixed-precision form of the mechanical solve: an FP64 flexible conjugate
// gradient (FCG(1), one-term orthogonalisation) on the finest level whose
// preconditioner is one FP32 multigrid V-cycle.  Every preconditioner kernel
// is templated on its arithmetic type P: P = float is the mixed-precision
// form, P = double the same V-cycle in FP64 (the like-for-like comparison of
// the two precisions at the same tolerance).
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

// FOLD: cubic grid, the grid scale w_g/h^2 folded into the moduli (the FP32
// finest level's Gauss-point moduli carry it from the set-up, the FP64
// finest level scales nodal K and G as its planes are committed, coarse
// levels scale K_e, G_e per element); see mech_element.
template <class P, int MODE, bool FINE, bool FOLD = false, bool LM = false>
struct MechOpP {
  using T = P;
  using TR = P;
  // the finest level reads nodal K, G, d and the split flag, as the FAS
  // passes do (LM: the precomputed FP32 Gauss-point moduli instead)
  static constexpr bool F64 = std::is_same<P, double>::value;
  static_assert(!(LM && F64), "the precomputed moduli are FP32");
  static constexpr bool NODAL = FINE && !LM;
  static constexpr int NC = 3;
  static constexpr int NU = 3;
  static constexpr int NV = NODAL ? 6 : NU;
  static constexpr int NR = NV;
  static constexpr bool NEEDS_OWN = (MODE != F_RES);
  static constexpr int NSL = F64 ? NSLOT : HPCR_MIXF_SLOTS, MIN_WAVES = F64 ? 1 : HPCR_MIXF_WAVES;
  using Raw = P[NR];
  struct Acc {
    double a[2];
    __device__ Acc() { a[0] = a[1] = 0.0; }
  };
  struct State {};
  // finest level, FP32: four float4 per element, (lambda, mu) of Gauss
  // points 2k and 2k+1 in float4 k (16-byte loads and LDS accesses);
  // finest level, FP64: the split flag; coarse levels: K_e, G_e
  using TE = std::conditional_t<FINE, std::conditional_t<NODAL, int, float4>, P>;
  static constexpr int NE = FINE ? (NODAL ? 1 : 4) : 2;
  MixCoefT<P> C;
  const P* s;            // stencil vector
  P* dout;               // CHEB: new direction; POWER: D^-1 A s
  P* r;                  // CHEB, RES: residual, updated in place
  P* x;                  // CHEB: iterate, updated in place
  const P* invD;
  P cd, cr;              // CHEB: d = cd s + cr D^-1 r
  const double* qprev;   // CHEB (optional): accumulate x.qprev into S_ZQ
  P sx, sy, sz, tx_, ty_, tz_;
  P fs;                  // FOLD: w_g / h^2
  Reduce R;
  size_t N;

  __device__ __forceinline__ State state() const { return State{}; }

  // node values through buffer resources: a node outside the grid (or past
  // the chunk) reads zeros without a branch
  __device__ __forceinline__ void fetch(const LevelGeom& L, int gx, int gy, int gz, bool ok, const State&,
                                        P (&v)[NR]) const {
    const unsigned ob = ok ? nid32(L, gx, gy, gz) * (unsigned)sizeof(P) : BOFF_NONE;
    const size_t nb = N * sizeof(P);
#pragma unroll
    for (int c = 0; c < 3; c++) v[c] = bldp<P>(brsrc(s + c * N, nb), ob);
    if constexpr (NODAL) {
      v[3] = bldp<P>(brsrc(C.Kn, nb), ob);
      v[4] = bldp<P>(brsrc(C.Gn, nb), ob);
      v[5] = bldp<P>(brsrc(C.dn, nb), ob);
    }
  }
  __device__ __forceinline__ void efetch(const LevelGeom& L, int ex, int ey, int ez, bool ok, TE (&e)[NE]) const {
    if constexpr (NODAL) {
      e[0] = (int)bld_i8(brsrc(C.m.flag, N), ok ? nid32(L, ex, ey, ez) : BOFF_NONE);
    } else {
      efetch_lm_or_coarse(L, ex, ey, ez, ok, e);
    }
  }
  __device__ __forceinline__ void efetch_lm_or_coarse(const LevelGeom& L, int ex, int ey, int ez, bool ok,
                                                      TE (&e)[NE]) const {
#pragma unroll
    for (int k = 0; k < NE; k++) e[k] = TE{};
    if (!ok) return;
    const unsigned q = nid32(L, ex, ey, ez);
    if constexpr (FINE) {
      // 64 bytes per element: the byte offset of a level's elements fits in
      // 32 bits up to 2^26 elements; beyond, index in 64 bits
      const float4* p = reinterpret_cast<const float4*>(C.lm) + 4 * (size_t)q;
#pragma unroll
      for (int k = 0; k < 4; k++) e[k] = p[k];
    } else {
      e[0] = C.Ke[q];
      e[1] = C.Ge[q];
    }
  }
  __device__ __forceinline__ void commit(const P (&v)[NR], const State&, P (&o)[NV]) const {
#pragma unroll
    for (int k = 0; k < NV; k++) o[k] = v[k];
    if constexpr (NODAL && FOLD) {
      o[3] *= fs;
      o[4] *= fs;
    }
  }

  __device__ __forceinline__ void element(const LevelGeom& L, const P (&sN)[NSL][NV][PLANE], int s0, int tx,
                                          int ty, int ex, int ey, int ez, const State&, const TE (&e)[NE],
                                          P (&f)[3][2][2][2]) const {
    P u[3][2][2][2];
#pragma unroll
    for (int c = 0; c < 3; c++) gather8(sN, s0, tx, ty, c, u[c]);
    if constexpr (NODAL) {
      P Kn[2][2][2], Gn[2][2][2], dn[2][2][2], lam[2][2][2], mu[2][2][2];
      gather8(sN, s0, tx, ty, 3, Kn);
      gather8(sN, s0, tx, ty, 4, Gn);
      gather8(sN, s0, tx, ty, 5, dn);
      fine_lame(Kn, Gn, dn, e[0], (P)C.m.small_k, lam, mu);
      mech_element<false, !FOLD>(u, lam, mu, P(0), P(0), sx, sy, sz, tx_, ty_, tz_, f);
    } else if constexpr (FINE) {
      float lam[2][2][2], mu[2][2][2];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int g0 = 2 * k, g1 = 2 * k + 1;
        lam[g0 >> 2][(g0 >> 1) & 1][g0 & 1] = e[k].x;
        mu[g0 >> 2][(g0 >> 1) & 1][g0 & 1] = e[k].y;
        lam[g1 >> 2][(g1 >> 1) & 1][g1 & 1] = e[k].z;
        mu[g1 >> 2][(g1 >> 1) & 1][g1 & 1] = e[k].w;
      }
      mech_element<false, !FOLD>(u, lam, mu, 0.0f, 0.0f, sx, sy, sz, tx_, ty_, tz_, f);
    } else {
      P lam0 = e[0] - (P(2) / P(3)) * e[1], mu0 = e[1];
      if constexpr (FOLD) {
        lam0 *= fs;
        mu0 *= fs;
      }
      P dl[2][2][2], dm[2][2][2];
      mech_element<true, !FOLD>(u, dl, dm, lam0, mu0, sx, sy, sz, tx_, ty_, tz_, f);
    }
  }

  // finest level: each element takes its bottom node plane's interpolated
  // moduli and in-plane gradient parts from the element below
  using Carry = std::conditional_t<NODAL && kDelta<P>, FineCarry<P>, void>;
  template <class C_>
  __device__ __forceinline__ void element_planes(const LevelGeom& L, const P (&sN)[NSL][NV][PLANE], int s0, int tx,
                                                 int ty, int ex, int ey, int ez, const State&, const TE (&e)[NE],
                                                 P (&f)[3][2][2][2], C_& cr, bool first) const {
    P u[3][2][2][2], Kn[2][2][2], Gn[2][2][2], dn[2][2][2], lam[2][2][2], mu[2][2][2];
#pragma unroll
    for (int c = 0; c < 3; c++) gather8(sN, s0, tx, ty, c, u[c]);
    gather8(sN, s0, tx, ty, 3, Kn);
    gather8(sN, s0, tx, ty, 4, Gn);
    gather8(sN, s0, tx, ty, 5, dn);
    fine_lame_planes(Kn, Gn, dn, e[0], (P)C.m.small_k, cr, first, lam, mu);
    mech_element_planes<!FOLD>(u, lam, mu, cr, first, sx, sy, sz, tx_, ty_, tz_, f);
  }

  // the epilogue's inputs at the owner's node, read one plane ahead
  static constexpr bool CHEB = (MODE == F_CHEB || MODE == F_CHEBQ);
  struct NData {
    P r[3], x[3], iD[3];
    double qp[MODE == F_CHEBQ ? 3 : 1];
  };
  __device__ __forceinline__ void nfetch(const LevelGeom& L, int i, int j, int k, NData& n) const {
    const unsigned q = nid32(L, i, j, k), ob = q * (unsigned)sizeof(P);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      if constexpr (MODE != F_POWER) n.r[c] = ldo(r + c * N, ob);
      if constexpr (CHEB) n.x[c] = ldo(x + c * N, ob);
      if constexpr (MODE != F_RES) n.iD[c] = ldo(invD + c * N, ob);
      if constexpr (MODE == F_CHEBQ) n.qp[c] = ldo(qprev + c * N, 8u * q);
    }
  }

  __device__ __forceinline__ void epilogue(const LevelGeom& L, const P (&own)[3], int i, int j, int k, int dm,
                                           const P (&v)[3], const State&, const NData& nd,
                                           Acc& acc) const {
    const unsigned ob = nid32(L, i, j, k) * (unsigned)sizeof(P);
    if constexpr (CHEB) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const P rn = ((dm >> c) & 1) ? P(0) : nd.r[c] - v[c];
        const P dn = cd * own[c] + cr * nd.iD[c] * rn;
        const P xn = nd.x[c] + dn;
        sto(r + c * N, ob, rn);
        sto(dout + c * N, ob, dn);
        sto(x + c * N, ob, xn);
        if constexpr (MODE == F_CHEBQ) acc.a[0] += (double)xn * nd.qp[c];
      }
    } else if constexpr (MODE == F_RES) {
#pragma unroll
      for (int c = 0; c < 3; c++) sto(r + c * N, ob, ((dm >> c) & 1) ? P(0) : nd.r[c] - v[c]);
    } else {  // F_POWER
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const bool dir = (dm >> c) & 1;
        const P di = nd.iD[c];
        sto(dout + c * N, ob, dir ? P(0) : di * v[c]);
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

// ---------------------------------------------------------------------------
// The FP32 finest-level element contraction on the matrix cores
// (v_mfma_f32_16x16x4_f32), the reference's B^T D B element product
// (mechanical_pcg.cpp:114-172) as batched GEMMs.  A wave holds 64 elements
// (tx = lane % 32, ty = 2 w + lane / 32), in four groups of 16.
//   1. gradient: for every component c and group, D (16 x 16) = A (16 x 8) U
//      (8 x 16) in two K-steps; row 4q + r of A is the derivative along axis q
//      at Gauss-point pair r (its sum-factorised B-matrix row: an edge
//      difference times two 1-D weights), rows 12..15 are zero, column n is
//      the group's element n, U holds the element's 8 nodal values (read from
//      the LDS node planes in the operand layout);
//   2. each lane reads its own element's 36 derivatives from a per-wave LDS
//      buffer, forms the stresses at the 8 Gauss points with its (lambda, mu)
//      and writes back the 36 transposed-contraction inputs (Sx, Sy, Sz);
//      the (lambda, mu) loads are issued from global memory at the start, so
//      the gradient products cover their latency, and they take no LDS (the
//      LDS goes to the wave buffers: three blocks per CU);
//   3. transposed contraction: F (16 x 16) = A^T (16 x 12) S (12 x 16) in three
//      K-steps; rows 0..7 are the element's 8 nodal forces;
//   4. each lane reads its own element's 24 forces.
// FP32 products with FP32 accumulation.  Operand layout of the instruction
// (one wave): A[i][k] in lane 16 k + i, B[k][j] in lane 16 k + j, D[4 q + r][j]
// in register r of lane 16 q + j.  The LDS buffer is private to the wave; its
// lanes meet only through in-order LDS accesses of that wave.
// ---------------------------------------------------------------------------
typedef float mf4 __attribute__((ext_vector_type(4)));
constexpr int MF_STRIDE = 36;  // floats per element in the wave's LDS buffer

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// row m = 4 q + r of the gradient matrix at node k (bits iz, iy, ix), with the
// axis scales s[q]
__device__ __forceinline__ float mf_grad_coef(int m, int k, const float (&s)[3]) {
  const int q = m >> 2, r = m & 3;
  if (q > 2) return 0.0f;
  const int iz = k >> 2, iy = (k >> 1) & 1, ix = k & 1;
  const int a = r >> 1, b = r & 1;
  auto Lf = [](int i, int g) { return (i == g) ? (float)GA : (float)GB; };
  float v;
  if (q == 0)
    v = (ix ? 1.0f : -1.0f) * Lf(iy, b) * Lf(iz, a);  // d/dx at (gz, gy) = (a, b)
  else if (q == 1)
    v = (iy ? 1.0f : -1.0f) * Lf(ix, b) * Lf(iz, a);  // d/dy at (gz, gx)
  else
    v = (iz ? 1.0f : -1.0f) * Lf(ix, b) * Lf(iy, a);  // d/dz at (gy, gx)
  return v * s[q];
}

template <int MODE, bool FOLD>
struct MechOpMfma : MechOpP<float, MODE, true, FOLD, true> {
  using Base = MechOpP<float, MODE, true, FOLD, true>;
  // no per-element data through the LDS plane pipeline (element() loads it)
  using TE = float;
  static constexpr int NE = 1;
  static constexpr int NSL = Base::NSL, NV = Base::NV;

  __device__ __forceinline__ void efetch(const LevelGeom&, int, int, int, bool, TE (&e)[NE]) const { e[0] = 0.0f; }

  __device__ __forceinline__ void element(const LevelGeom& L, const float (&sN)[NSL][NV][PLANE], int s0, int tx,
                                          int ty, int ex, int ey, int ez, const typename Base::State&,
                                          const TE (&)[NE], float (&f)[3][2][2][2]) const {
    __shared__ float sM[TPB / 64][64 * MF_STRIDE];
    // this lane's Gauss-point moduli, zero for an element outside the grid
    float4 lmv[4];
    {
      const bool ok = ex >= 0 && ey >= 0 && ex <= L.nx - 2 && ey <= L.ny - 2 && ez >= 0 && ez <= L.nz - 2;
      const float4* p = reinterpret_cast<const float4*>(this->C.lm) + 4 * nidx(L, ok ? ex : 0, ok ? ey : 0, ok ? ez : 0);
#pragma unroll
      for (int k = 0; k < 4; k++) lmv[k] = ok ? p[k] : float4{0.0f, 0.0f, 0.0f, 0.0f};
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    float* M = sM[w];
    const int i16 = lane & 15, k4 = lane >> 4;
    const int s1 = (NSL == 3) ? (s0 == 2 ? 0 : s0 + 1) : (s0 ^ 1);
    const float sg[3] = {FOLD ? 1.0f : this->sx, FOLD ? 1.0f : this->sy, FOLD ? 1.0f : this->sz};
    const float st[3] = {FOLD ? 1.0f : this->tx_, FOLD ? 1.0f : this->ty_, FOLD ? 1.0f : this->tz_};
    // operands of A (gradient, K-steps 0..1) and A^T (transposed, K-steps 0..2)
    const float ag0 = mf_grad_coef(i16, k4, sg), ag1 = mf_grad_coef(i16, 4 + k4, sg);
    float at[3];
#pragma unroll
    for (int q = 0; q < 3; q++) at[q] = (i16 < 8) ? mf_grad_coef(4 * q + k4, i16, st) : 0.0f;
    // 1. gradients of the wave's 64 elements
    const int iy = k4 >> 1, ix = k4 & 1;
#pragma unroll
    for (int g = 0; g < 4; g++) {
      const int el = 16 * g + i16;                        // element of this operand column
      const int qe = (el & 31) + PX * (2 * w + (el >> 5)) + ix + PX * iy;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        mf4 d = {0.0f, 0.0f, 0.0f, 0.0f};
        d = __builtin_amdgcn_mfma_f32_16x16x4f32(ag0, sN[s0][c][qe], d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_16x16x4f32(ag1, sN[s1][c][qe], d, 0, 0, 0);
        if (k4 < 3) *reinterpret_cast<mf4*>(&M[el * MF_STRIDE + c * 12 + 4 * k4]) = d;
      }
    }
    wave_lds_sync();
    // 2. stresses at the Gauss points of this lane's element
    {
      float G[3][3][4];
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
        for (int q = 0; q < 3; q++) {
          const mf4 v = *reinterpret_cast<const mf4*>(&M[lane * MF_STRIDE + c * 12 + 4 * q]);
          G[c][q][0] = v[0];
          G[c][q][1] = v[1];
          G[c][q][2] = v[2];
          G[c][q][3] = v[3];
        }
      float lam[2][2][2], mu[2][2][2];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int g0 = 2 * k, g1 = 2 * k + 1;
        lam[g0 >> 2][(g0 >> 1) & 1][g0 & 1] = lmv[k].x;
        mu[g0 >> 2][(g0 >> 1) & 1][g0 & 1] = lmv[k].y;
        lam[g1 >> 2][(g1 >> 1) & 1][g1 & 1] = lmv[k].z;
        mu[g1 >> 2][(g1 >> 1) & 1][g1 & 1] = lmv[k].w;
      }
      float Sx[3][4], Sy[3][4], Sz[3][4];
#pragma unroll
      for (int gz = 0; gz < 2; gz++)
#pragma unroll
        for (int gy = 0; gy < 2; gy++)
#pragma unroll
          for (int gx = 0; gx < 2; gx++) {
            const int px = 2 * gz + gy, py = 2 * gz + gx, pz = 2 * gy + gx;
            const float exx = G[0][0][px], eyy = G[1][1][py], ezz = G[2][2][pz];
            const float gxy = G[1][0][px] + G[0][1][py];
            const float gxz = G[2][0][px] + G[0][2][pz];
            const float gyz = G[2][1][py] + G[1][2][pz];
            const float l = lam[gz][gy][gx], m = mu[gz][gy][gx];
            const float lt = l * (exx + eyy + ezz), m2 = 2.0f * m;
            const float sxx = lt + m2 * exx, syy = lt + m2 * eyy, szz = lt + m2 * ezz;
            const float sxy = m * gxy, sxz = m * gxz, syz = m * gyz;
            auto put = [](float& a, float v, bool first) {
              if (first) a = v; else a += v;
            };
            put(Sx[0][px], sxx, gx == 0); put(Sx[1][px], sxy, gx == 0); put(Sx[2][px], sxz, gx == 0);
            put(Sy[0][py], sxy, gy == 0); put(Sy[1][py], syy, gy == 0); put(Sy[2][py], syz, gy == 0);
            put(Sz[0][pz], sxz, gz == 0); put(Sz[1][pz], syz, gz == 0); put(Sz[2][pz], szz, gz == 0);
          }
#pragma unroll
      for (int c = 0; c < 3; c++) {
        float* o = &M[lane * MF_STRIDE + c * 12];
        *reinterpret_cast<mf4*>(o + 0) = mf4{Sx[c][0], Sx[c][1], Sx[c][2], Sx[c][3]};
        *reinterpret_cast<mf4*>(o + 4) = mf4{Sy[c][0], Sy[c][1], Sy[c][2], Sy[c][3]};
        *reinterpret_cast<mf4*>(o + 8) = mf4{Sz[c][0], Sz[c][1], Sz[c][2], Sz[c][3]};
      }
    }
    wave_lds_sync();
    // 3. nodal forces: component c's inputs are read for all four groups
    //    before its forces overwrite them
#pragma unroll
    for (int c = 0; c < 3; c++) {
      mf4 d[4];
#pragma unroll
      for (int g = 0; g < 4; g++) {
        const int el = 16 * g + i16;
        d[g] = mf4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int q = 0; q < 3; q++)
          d[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(at[q], M[el * MF_STRIDE + c * 12 + 4 * q + k4], d[g], 0, 0, 0);
      }
      wave_lds_sync();
#pragma unroll
      for (int g = 0; g < 4; g++)
        if (k4 < 2) *reinterpret_cast<mf4*>(&M[(16 * g + i16) * MF_STRIDE + c * 12 + 4 * k4]) = d[g];
    }
    wave_lds_sync();
    // 4. this lane's element: f[c][iz][iy][ix] = F[c][4 iz + 2 iy + ix]
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const mf4 a = *reinterpret_cast<const mf4*>(&M[lane * MF_STRIDE + c * 12 + 0]);
      const mf4 b = *reinterpret_cast<const mf4*>(&M[lane * MF_STRIDE + c * 12 + 4]);
      f[c][0][0][0] = a[0];
      f[c][0][0][1] = a[1];
      f[c][0][1][0] = a[2];
      f[c][0][1][1] = a[3];
      f[c][1][0][0] = b[0];
      f[c][1][0][1] = b[1];
      f[c][1][1][0] = b[2];
      f[c][1][1][1] = b[3];
    }
    // the buffer is rewritten by the next layer's gradients only after
    // every lane has read its forces
    wave_lds_sync();
  }
};

// FCG pass A (FP64, finest level): p = z + beta p_prev with z the
// preconditioner output (type P); q = A p; p.q, p.r -> alpha.  LAG: also
// u += alpha_prev p_prev at the own nodes, the update of the previous
// iteration that its lean pass B left out (the same fma, one kernel later).
template <class P, bool LAG, bool FOLD = false>
struct FcgOpA {
  using T = double;
  using TR = double;
  static constexpr int NC = 3;
  static constexpr int NU = 3;
  static constexpr int NV = 6;
  static constexpr int NR = 9;
  static constexpr bool NEEDS_OWN = true;
  static constexpr int NSL = NSLOT, MIN_WAVES = 1;
  // raw node values in flight: z stays in its own type until the commit, so
  // that no conversion consumes a load before the layer that needs it
  struct Raw {
    P z[3];
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
  const P* z;
  const double* p_old;
  double* p_new;
  double* q;
  const double* r;
  double* u;  // LAG
  int first;
  double sx, sy, sz, tx_, ty_, tz_;
  double fs;  // FOLD: w_g / h^2
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
      v.z[c] = P(0);
      v.p[c] = 0.0;
    }
    v.K = v.G = v.d = 0.0;
    if (!ok) return;
    const unsigned q = nid32(L, gx, gy, gz), ob = 8u * q;
#pragma unroll
    for (int c = 0; c < 3; c++) v.z[c] = ldo(z + c * N, q * (unsigned)sizeof(P));
    if (!first) {
#pragma unroll
      for (int c = 0; c < 3; c++) v.p[c] = ldo(p_old + c * N, ob);
    }
    v.K = ldo(C.K, ob);
    v.G = ldo(C.G, ob);
    v.d = ldo(C.d, ob);
  }
  __device__ __forceinline__ void efetch(const LevelGeom& L, int ex, int ey, int ez, bool ok, int (&e)[1]) const {
    e[0] = ok ? (int)ldo(C.flag, nid32(L, ex, ey, ez)) : 0;
  }
  __device__ __forceinline__ void commit(const Raw& v, const State& st, double (&w)[NV]) const {
#pragma unroll
    for (int c = 0; c < 3; c++) w[c] = (double)v.z[c] + st.beta * v.p[c];
    w[3] = FOLD ? v.K * fs : v.K;
    w[4] = FOLD ? v.G * fs : v.G;
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
    mech_element<false, !FOLD>(u, lam, mu, 0.0, 0.0, sx, sy, sz, tx_, ty_, tz_, f);
  }
  // each element takes its bottom node plane's interpolated moduli from the
  // element below (the gradient parts stay per element: carrying them too
  // leaves this pass one wave per SIMD)
  using Carry = std::conditional_t<kDelta<double>, FineCarry<double>, void>;
  template <class C_>
  __device__ __forceinline__ void element_planes(const LevelGeom& L, const double (&sN)[NSL][NV][PLANE], int s0,
                                                 int tx, int ty, int ex, int ey, int ez, const State&,
                                                 const int (&e)[1], double (&f)[3][2][2][2], C_& cr,
                                                 bool first) const {
    double u[3][2][2][2], Kn[2][2][2], Gn[2][2][2], dn[2][2][2], lam[2][2][2], mu[2][2][2];
#pragma unroll
    for (int c = 0; c < 3; c++) gather8(sN, s0, tx, ty, c, u[c]);
    gather8(sN, s0, tx, ty, 3, Kn);
    gather8(sN, s0, tx, ty, 4, Gn);
    gather8(sN, s0, tx, ty, 5, dn);
    fine_lame_planes(Kn, Gn, dn, e[0], C.small_k, cr, first, lam, mu);
    mech_element<false, !FOLD>(u, lam, mu, 0.0, 0.0, sx, sy, sz, tx_, ty_, tz_, f);
  }
  struct NData {
    double r[3];
    double u[LAG ? 3 : 1], p[LAG ? 3 : 1];
  };
  __device__ __forceinline__ void nfetch(const LevelGeom& L, int i, int j, int k, NData& n) const {
    const unsigned ob = 8u * nid32(L, i, j, k);
#pragma unroll
    for (int c = 0; c < 3; c++) n.r[c] = ldo(r + c * N, ob);
    if constexpr (LAG) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        n.u[c] = ldo(u + c * N, ob);
        n.p[c] = ldo(p_old + c * N, ob);
      }
    }
  }
  __device__ __forceinline__ void epilogue(const LevelGeom& L, const double (&own)[3], int i, int j, int k, int dm,
                                           const double (&v)[3], const State& st, const NData& nd,
                                           Acc& acc) const {
    const unsigned ob = 8u * nid32(L, i, j, k);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      sto(p_new + c * N, ob, own[c]);
      sto(q + c * N, ob, v[c]);
      acc.a[0] += own[c] * v[c];
      acc.a[1] += own[c] * nd.r[c];
      if constexpr (LAG) sto(u + c * N, ob, fma(st.aprev, nd.p[c], nd.u[c]));
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

template <class P, int MODE, bool FINE, bool FOLD, bool MFMA = false>
static void run_mixp(hipStream_t st, const LevelGeom& L, const MixCoefT<P>& C, const P* s, P* dout, P* r, P* x,
                     const P* invD, P cd, P cr, const double* qprev, const Reduce& R) {
  using Op = std::conditional_t<MFMA, MechOpMfma<MODE, FOLD>, MechOpP<P, MODE, FINE, FOLD>>;
  Op op{};
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
  op.sx = (P)(1.0 / L.hx);
  op.sy = (P)(1.0 / L.hy);
  op.sz = (P)(1.0 / L.hz);
  op.tx_ = (P)(wg / L.hx);
  op.ty_ = (P)(wg / L.hy);
  op.tz_ = (P)(wg / L.hz);
  op.fs = (P)(wg / (L.hx * L.hx));
  op.R = R;
  op.N = (size_t)L.nx * L.ny * L.nz;
  if (record<TileK<Op>>(mix_tile_grid(L), {L, op})) return;
  tile_kernel<Op><<<mix_tile_grid(L), dim3(TPB), 0, st>>>(L, op);
}

template <class P, int MODE>
static void run_mixp_level(hipStream_t st, const LevelGeom& L, const MixCoefT<P>& C, const P* s, P* dout, P* r,
                           P* x, const P* invD, P cd, P cr, const double* qprev, const Reduce& R) {
  // the FP32 finest level's moduli carry the grid scale when the set-up
  // folded it (mix_lame with the cubic grid's w_g/h^2)
  const bool fold = cubic_grid(L);
  if (C.fine) {
    if constexpr (std::is_same<P, float>::value) {
      if (C.mfma) {
        if (fold)
          run_mixp<P, MODE, true, true, true>(st, L, C, s, dout, r, x, invD, cd, cr, qprev, R);
        else
          run_mixp<P, MODE, true, false, true>(st, L, C, s, dout, r, x, invD, cd, cr, qprev, R);
        return;
      }
    }
    if (fold)
      run_mixp<P, MODE, true, true>(st, L, C, s, dout, r, x, invD, cd, cr, qprev, R);
    else
      run_mixp<P, MODE, true, false>(st, L, C, s, dout, r, x, invD, cd, cr, qprev, R);
  } else {
    if (fold)
      run_mixp<P, MODE, false, true>(st, L, C, s, dout, r, x, invD, cd, cr, qprev, R);
    else
      run_mixp<P, MODE, false, false>(st, L, C, s, dout, r, x, invD, cd, cr, qprev, R);
  }
}

template <class P>
void launch_mix_cheb(hipStream_t st, const LevelGeom& L, const MixCoefT<P>& C, const P* s, P* dout, P* r, P* x,
                     const P* invD, P cd, P cr, const double* qprev, const Reduce& R) {
  if (qprev)
    run_mixp_level<P, F_CHEBQ>(st, L, C, s, dout, r, x, invD, cd, cr, qprev, R);
  else
    run_mixp_level<P, F_CHEB>(st, L, C, s, dout, r, x, invD, cd, cr, nullptr, R);
}
template <class P>
void launch_mix_res(hipStream_t st, const LevelGeom& L, const MixCoefT<P>& C, const P* s, P* r) {
  Reduce R{};
  run_mixp_level<P, F_RES>(st, L, C, s, nullptr, r, nullptr, nullptr, P(0), P(0), nullptr, R);
}
template <class P>
void launch_mix_power(hipStream_t st, const LevelGeom& L, const MixCoefT<P>& C, const P* v, P* w, const P* invD,
                      const Reduce& R) {
  run_mixp_level<P, F_POWER>(st, L, C, v, w, nullptr, nullptr, invD, P(0), P(0), nullptr, R);
}

template <class P, bool LAG, bool FOLD>
static void run_fcg_a(hipStream_t st, const LevelGeom& L, const MechCoef& C, const P* z, const double* p_old,
                      double* p_new, double* q, const double* r, int first, double* u, const Reduce& R) {
  FcgOpA<P, LAG, FOLD> op{};
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
  op.fs = wg / (L.hx * L.hx);
  op.R = R;
  op.N = (size_t)L.nx * L.ny * L.nz;
  using Op = FcgOpA<P, LAG, FOLD>;
  if (record<TileK<Op>>(mix_tile_grid(L), {L, op})) return;
  tile_kernel<Op><<<mix_tile_grid(L), dim3(TPB), 0, st>>>(L, op);
}
template <class P>
void launch_fcg_a(hipStream_t st, const LevelGeom& L, const MechCoef& C, const P* z, const double* p_old,
                  double* p_new, double* q, const double* r, int first, double* u_lag, const Reduce& R) {
  const bool fold = cubic_grid(L);
  if (u_lag && !first) {
    if (fold)
      run_fcg_a<P, true, true>(st, L, C, z, p_old, p_new, q, r, 0, u_lag, R);
    else
      run_fcg_a<P, true, false>(st, L, C, z, p_old, p_new, q, r, 0, u_lag, R);
  } else {
    if (fold)
      run_fcg_a<P, false, true>(st, L, C, z, p_old, p_new, q, r, first, nullptr, R);
    else
      run_fcg_a<P, false, false>(st, L, C, z, p_old, p_new, q, r, first, nullptr, R);
  }
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
template <class P>
__global__ __launch_bounds__(MX_TPB) void mix_start_kernel(unsigned N, const double* __restrict__ r64,
                                                           P* __restrict__ r32, P* __restrict__ d,
                                                           P* __restrict__ x, const P* __restrict__ invD, P cr0) {
  for (unsigned q = blockIdx.x * MX_TPB + threadIdx.x; q < N; q += gridDim.x * MX_TPB) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const size_t o = (size_t)c * N + q;
      const P rv = (P)r64[o];
      const P dv = cr0 * invD[o] * rv;
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
template <class P, bool LDSF, bool LEAN>
__device__ __forceinline__ void fcg_b_body(const LevelGeom& L, double* __restrict__ u, double* __restrict__ r,
                                           const double* __restrict__ p, const double* __restrict__ q,
                                           P* __restrict__ r32, P* __restrict__ d, P* __restrict__ x,
                                           const P* __restrict__ invD, P cr0, const Reduce& R, const Blk& bk) {
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
      const P rf = (P)rv;
      const P dv = cr0 * invD[o] * rf;
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
template <class P, bool LDSF, bool LEAN>
__global__ __launch_bounds__(MX_TPB) void fcg_b_kernel(const LevelGeom L, double* __restrict__ u,
                                                       double* __restrict__ r, const double* __restrict__ p,
                                                       const double* __restrict__ q, P* __restrict__ r32,
                                                       P* __restrict__ d, P* __restrict__ x,
                                                       const P* __restrict__ invD, P cr0, const Reduce R) {
  fcg_b_body<P, LDSF, LEAN>(L, u, r, p, q, r32, d, x, invD, cr0, R, Blk::hw());
}
template <class P>
struct FcgBArgs {
  LevelGeom L;
  double* u;
  double* r;
  const double* p;
  const double* q;
  P* r32;
  P* d;
  P* x;
  const P* invD;
  P cr0;
  Reduce R;
};
template <class P, bool LDSF, bool LEAN>
struct FcgBK {
  using Args = FcgBArgs<P>;
  static constexpr int NT = MX_TPB, MINW = 1;
  static dim3 block() { return dim3(MX_TPB); }
  __device__ __forceinline__ static void run(const Args& a, const Blk& bk) {
    fcg_b_body<P, LDSF, LEAN>(a.L, a.u, a.r, a.p, a.q, a.r32, a.d, a.x, a.invD, a.cr0, a.R, bk);
  }
};

__global__ __launch_bounds__(MX_TPB) void fcg_flush_kernel(size_t n, double* __restrict__ u,
                                                           const double* __restrict__ p,
                                                           const double* __restrict__ scal) {
  const double alpha = scal[S_APREV];
  for (size_t o = (size_t)blockIdx.x * MX_TPB + threadIdx.x; o < n; o += (size_t)gridDim.x * MX_TPB)
    u[o] = fma(alpha, p[o], u[o]);
}

// restriction of the preconditioner's residual (full weighting, as the FP64
// transfer of the FAS path);
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

template <class P>
__device__ __forceinline__ void mix_restrict_body(const LevelGeom& Lc, const LevelGeom& Lf,
                                                  const P* __restrict__ rf, P* __restrict__ rc, P* __restrict__ dc,
                                                  P* __restrict__ xc, const P* __restrict__ invD, P cr0, int start,
                                                  const Blk& bk) {
  const unsigned qn = bk.x * 256 + threadIdx.x;
  if (qn >= (unsigned)Nof(Lc)) return;
  int i, j, k;
  mix_ijk(Lc, qn, i, j, k);
  const size_t Nc = Nof(Lc), Nf = Nof(Lf);
  const Restrict27<P> R(Lf, i, j, k);
  const size_t qc = nidx(Lc, i, j, k);
  const int dm = dir_mask(Lc, i, j, k);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    P acc = R(rf + c * Nf);
    if ((dm >> c) & 1) acc = P(0);
    const size_t o = c * Nc + qc;
    rc[o] = acc;
    if (start) {
      const P dv = cr0 * invD[o] * acc;
      dc[o] = dv;
      xc[o] = dv;
    } else {
      xc[o] = P(0);
    }
  }
}
template <class P>
__global__ __launch_bounds__(256) void mix_restrict_kernel(const LevelGeom Lc, const LevelGeom Lf,
                                                           const P* __restrict__ rf, P* __restrict__ rc,
                                                           P* __restrict__ dc, P* __restrict__ xc,
                                                           const P* __restrict__ invD, P cr0, int start) {
  mix_restrict_body<P>(Lc, Lf, rf, rc, dc, xc, invD, cr0, start, Blk::hw());
}
template <class P>
struct MixRestrictK {
  struct Args {
    LevelGeom Lc;
    LevelGeom Lf;
    const P* rf;
    P* rc;
    P* dc;
    P* xc;
    const P* invD;
    P cr0;
    int start;
  };
  static constexpr int NT = 256, MINW = 1;
  static dim3 block() { return dim3(256); }
  __device__ __forceinline__ static void run(const Args& a, const Blk& bk) {
    mix_restrict_body<P>(a.Lc, a.Lf, a.rf, a.rc, a.dc, a.xc, a.invD, a.cr0, a.start, bk);
  }
};

// prolongation of the coarse correction: c = P e (zero on Dirichlet DOFs),
// x_f += c, c kept as the stencil vector of the first post-smoothing step.
// FP64: two x-adjacent fine nodes per thread (ProlongPair), which halves the
// coarse loads.  FP32: one node per thread (the pair's value for that node),
// since there the pairwise form's half-width stores cost more than the loads
// it saves (measured 160 -> 186 us per 257^3 prolongation).
template <class P>
constexpr bool kProlongPairs = std::is_same<P, double>::value;
template <class P>
__host__ __device__ __forceinline__ size_t mix_prolong_threads(const LevelGeom& Lf) {
  return kProlongPairs<P> ? prolong_pairs(Lf) : Nof(Lf);
}
template <class P>
__device__ __forceinline__ void mix_prolong_body(const LevelGeom& Lf, const LevelGeom& Lc, P* __restrict__ xf,
                                                 P* __restrict__ cf, const P* __restrict__ e, const Blk& bk) {
  const unsigned q = bk.x * 256 + threadIdx.x;
  if (q >= (unsigned)mix_prolong_threads<P>(Lf)) return;
  const size_t Nc = Nof(Lc), Nf = Nof(Lf);
  if constexpr (kProlongPairs<P>) {
    int ip, y, z;
    bool has1;
    pair_ijk(Lf, q, ip, has1, y, z);
    const ProlongPair<P> Pr(Lc, ip, has1, y, z);
    const int x = 2 * ip;
    const unsigned of = nid32(Lf, x, y, z) * (unsigned)sizeof(P), of1 = of + (unsigned)sizeof(P);
    const int dm0 = dir_mask(Lf, x, y, z);
    const int dm1 = has1 ? dir_mask(Lf, x + 1, y, z) : 0;
    P c0[3], c1[3];
#pragma unroll
    for (int c = 0; c < 3; c++) Pr(e + c * Nc, c0[c], c1[c]);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const P cv0 = ((dm0 >> c) & 1) ? P(0) : c0[c];
      sto(xf + c * Nf, of, ldo(xf + c * Nf, of) + cv0);
      sto(cf + c * Nf, of, cv0);
      if (has1) {
        const P cv1 = ((dm1 >> c) & 1) ? P(0) : c1[c];
        sto(xf + c * Nf, of1, ldo(xf + c * Nf, of1) + cv1);
        sto(cf + c * Nf, of1, cv1);
      }
    }
  } else {
    int x, y, z;
    mix_ijk(Lf, q, x, y, z);
    // an odd x is the odd node of pair x >> 1; an even x the even node of a
    // pair whose column-(i+1) loads read column i (has1 false)
    const bool odd = x & 1;
    const ProlongPair<P> Pr(Lc, x >> 1, odd, y, z);
    const unsigned of = nid32(Lf, x, y, z) * (unsigned)sizeof(P);
    const int dm = dir_mask(Lf, x, y, z);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      P c0, c1;
      Pr(e + c * Nc, c0, c1);
      const P cv = ((dm >> c) & 1) ? P(0) : (odd ? c1 : c0);
      sto(xf + c * Nf, of, ldo(xf + c * Nf, of) + cv);
      sto(cf + c * Nf, of, cv);
    }
  }
}
template <class P>
__global__ __launch_bounds__(256) void mix_prolong_kernel(const LevelGeom Lf, const LevelGeom Lc,
                                                          P* __restrict__ xf, P* __restrict__ cf,
                                                          const P* __restrict__ e) {
  mix_prolong_body<P>(Lf, Lc, xf, cf, e, Blk::hw());
}
template <class P>
struct MixProlongK {
  struct Args {
    LevelGeom Lf;
    LevelGeom Lc;
    P* xf;
    P* cf;
    const P* e;
  };
  static constexpr int NT = 256, MINW = 1;
  static dim3 block() { return dim3(256); }
  __device__ __forceinline__ static void run(const Args& a, const Blk& bk) {
    mix_prolong_body<P>(a.Lf, a.Lc, a.xf, a.cf, a.e, bk);
  }
};

// (lambda, mu) at the Gauss points of every finest-level element (the
// coefficients of MP_all_GAUSS_Point, computed in FP64 and stored as FP32)
// (times w_g/h^2 on a cubic grid, the scale the FP32 passes then leave out)
__global__ __launch_bounds__(256) void mix_lame_kernel(const LevelGeom L, const double* __restrict__ K,
                                                       const double* __restrict__ G, const double* __restrict__ d,
                                                       const signed char* __restrict__ flag, double small_k,
                                                       double scale, float* __restrict__ lm) {
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
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int c = 0; c < 2; c++) {
        lam[a][b][c] *= scale;
        mu[a][b][c] *= scale;
      }
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
  const double scale = cubic_grid(L) ? 0.125 * L.hx * L.hy * L.hz / (L.hx * L.hx) : 1.0;
  hipLaunchKernelGGL(mix_lame_kernel, g, dim3(64, 4), 0, st, L, K, G, d, flag, small_k, scale, lm);
}

// b = a, converted to b's type (a copy when the types agree)
template <class A, class B>
__device__ __forceinline__ void cvt_body(size_t n, const A* __restrict__ a, B* __restrict__ b, const Blk& bk) {
  for (size_t q = (size_t)bk.x * blockDim.x + threadIdx.x; q < n; q += (size_t)bk.gx * blockDim.x)
    b[q] = (B)a[q];
}
template <class A, class B>
__global__ void cvt_kernel(size_t n, const A* __restrict__ a, B* __restrict__ b) {
  cvt_body<A, B>(n, a, b, Blk::hw());
}
template <class A, class B>
struct CvtK {
  struct Args {
    size_t n;
    const A* a;
    B* b;
  };
  static constexpr int NT = 256, MINW = 1;
  static dim3 block() { return dim3(256); }
  __device__ __forceinline__ static void run(const Args& a, const Blk& bk) { cvt_body<A, B>(a.n, a.a, a.b, bk); }
};

// deterministic start vector of the power iteration, zero on Dirichlet DOFs
template <class P>
__global__ void power_init_kernel(const LevelGeom L, P* __restrict__ v) {
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
      v[(size_t)c * N + n] = ((dm >> c) & 1) ? P(0) : (P)val;
    }
  }
}

template <class P>
void launch_mix_start(hipStream_t st, const LevelGeom& L, const double* r64, P* r32, P* d, P* x, const P* invD,
                      P cr0) {
  const size_t N = Nof(L);
  hipLaunchKernelGGL(mix_start_kernel<P>, dim3(mx_blocks(N, 4096)), dim3(MX_TPB), 0, st, (unsigned)N, r64, r32, d, x,
                     invD, cr0);
}
template <class P>
void launch_fcg_b(hipStream_t st, const LevelGeom& L, double* u, double* r, const double* p, const double* q,
                  P* r32, P* d, P* x, const P* invD, P cr0, bool lean, const Reduce& R) {
  const dim3 g(passb_blocks(Nof(L))), b(MX_TPB);
  const bool ldsf = L.nx + L.ny + L.nz <= DIRF_MAX;
  const FcgBArgs<P> a{L, u, r, p, q, r32, d, x, invD, cr0, R};
  bool rec;
  if (ldsf && lean)
    rec = record<FcgBK<P, true, true>>(g, a);
  else if (ldsf)
    rec = record<FcgBK<P, true, false>>(g, a);
  else if (lean)
    rec = record<FcgBK<P, false, true>>(g, a);
  else
    rec = record<FcgBK<P, false, false>>(g, a);
  if (rec) return;
  if (ldsf && lean)
    hipLaunchKernelGGL((fcg_b_kernel<P, true, true>), g, b, 0, st, L, u, r, p, q, r32, d, x, invD, cr0, R);
  else if (ldsf)
    hipLaunchKernelGGL((fcg_b_kernel<P, true, false>), g, b, 0, st, L, u, r, p, q, r32, d, x, invD, cr0, R);
  else if (lean)
    hipLaunchKernelGGL((fcg_b_kernel<P, false, true>), g, b, 0, st, L, u, r, p, q, r32, d, x, invD, cr0, R);
  else
    hipLaunchKernelGGL((fcg_b_kernel<P, false, false>), g, b, 0, st, L, u, r, p, q, r32, d, x, invD, cr0, R);
}
void launch_fcg_flush(hipStream_t st, const LevelGeom& L, double* u, const double* p, const Reduce& R) {
  const size_t n = 3 * Nof(L);
  hipLaunchKernelGGL(fcg_flush_kernel, dim3(mx_blocks(n, 4096)), dim3(MX_TPB), 0, st, n, u, p,
                     static_cast<const double*>(R.scal));
}
template <class P>
void launch_mix_restrict(hipStream_t st, const LevelGeom& Lc, const LevelGeom& Lf, const P* rf, P* rc, P* dc, P* xc,
                         const P* invD, P cr0, int start) {
  const dim3 g((unsigned)((Nof(Lc) + 255) / 256));
  if (record<MixRestrictK<P>>(g, {Lc, Lf, rf, rc, dc, xc, invD, cr0, start})) return;
  hipLaunchKernelGGL(mix_restrict_kernel<P>, g, dim3(256), 0, st, Lc, Lf, rf, rc, dc, xc, invD, cr0, start);
}
template <class P>
void launch_mix_prolong(hipStream_t st, const LevelGeom& Lf, const LevelGeom& Lc, P* xf, P* cf, const P* ec) {
  const dim3 g((unsigned)((mix_prolong_threads<P>(Lf) + 255) / 256));
  if (record<MixProlongK<P>>(g, {Lf, Lc, xf, cf, ec})) return;
  hipLaunchKernelGGL(mix_prolong_kernel<P>, g, dim3(256), 0, st, Lf, Lc, xf, cf, ec);
}
template <class A, class B>
void launch_convert(hipStream_t st, size_t n, const A* a, B* b) {
  if (record<CvtK<A, B>>(dim3(mx_blocks(n, 4096)), {n, a, b})) return;
  hipLaunchKernelGGL((cvt_kernel<A, B>), dim3(mx_blocks(n, 4096)), dim3(256), 0, st, n, a, b);
}
template <class P>
void launch_power_init(hipStream_t st, const LevelGeom& L, P* v) {
  hipLaunchKernelGGL(power_init_kernel<P>, dim3(mx_blocks(Nof(L), 4096)), dim3(256), 0, st, L, v);
}

// the two preconditioner precisions
#define HPCR_MIX_INSTANTIATE(P)                                                                                   \
  template void launch_mix_cheb<P>(hipStream_t, const LevelGeom&, const MixCoefT<P>&, const P*, P*, P*, P*,       \
                                   const P*, P, P, const double*, const Reduce&);                                \
  template void launch_mix_res<P>(hipStream_t, const LevelGeom&, const MixCoefT<P>&, const P*, P*);               \
  template void launch_mix_power<P>(hipStream_t, const LevelGeom&, const MixCoefT<P>&, const P*, P*, const P*,    \
                                    const Reduce&);                                                               \
  template void launch_power_init<P>(hipStream_t, const LevelGeom&, P*);                                          \
  template void launch_fcg_a<P>(hipStream_t, const LevelGeom&, const MechCoef&, const P*, const double*, double*, \
                                double*, const double*, int, double*, const Reduce&);                             \
  template void launch_fcg_b<P>(hipStream_t, const LevelGeom&, double*, double*, const double*, const double*,    \
                                P*, P*, P*, const P*, P, bool, const Reduce&);                                    \
  template void launch_mix_start<P>(hipStream_t, const LevelGeom&, const double*, P*, P*, P*, const P*, P);       \
  template void launch_mix_restrict<P>(hipStream_t, const LevelGeom&, const LevelGeom&, const P*, P*, P*, P*,     \
                                       const P*, P, int);                                                         \
  template void launch_mix_prolong<P>(hipStream_t, const LevelGeom&, const LevelGeom&, P*, P*, const P*);
HPCR_MIX_INSTANTIATE(float)
HPCR_MIX_INSTANTIATE(double)
#undef HPCR_MIX_INSTANTIATE
template void launch_convert<double, float>(hipStream_t, size_t, const double*, float*);
template void launch_convert<float, double>(hipStream_t, size_t, const float*, double*);
template void launch_convert<double, double>(hipStream_t, size_t, const double*, double*);

}  // namespace hpcr
