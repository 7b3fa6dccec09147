// HIP kernels for gfx950 (MI355X): matrix-free trilinear-hex operators of
// HPCrack's degraded elasticity and AT2 phase field, the fused Jacobi-PCG
// passes, multigrid transfers, coarse coefficients and Gauss-point updates.
//
// Operator formulation.  On a uniform voxel grid the gradient of a trilinear
// field at the 2x2x2 Gauss points is sum-factorised: d/dx at Gauss point
// (gx,gy,gz) does not depend on gx and equals the x-edge differences of the
// element interpolated bilinearly in y and z.  The transposed contraction
// B^T sigma is factorised the same way.  This needs ~1.2 kflop per element
// instead of ~8 kflop for the reference's per-Gauss-point loops
// (mechanical_pcg.cpp:114-172), and gives the same operator up to rounding.
//
// Tiling.  A block of 256 threads owns a 32x8 column of elements (one per
// thread) and marches up a chunk of z-layers.  Each step evaluates one
// element layer from two node planes held in LDS (u and, on the finest level,
// K, G, d) while the global loads of the plane after next are in flight.  The
// 8 corner contributions of every element are exchanged so that each node
// sums its 4 in-plane neighbours in a fixed order: the x-neighbour is the
// adjacent lane (DPP row shift), the y-neighbour goes through LDS.  The node
// plane below the layer is then complete and handed to the fused epilogue
// (residual, PCG update, diagonal ...); the plane above is carried in
// registers.  No atomics: results are deterministic.  Tiles overlap by one
// element so that the 31x7 interior node columns are owned by one block.
//
// Small levels.  A level of at most SMALL_NODES nodes and SMALL_ELEMS
// elements runs a whole Jacobi-PCG smoother call in one workgroup
// (small_pcg_kernel): vectors in registers, p and the element contributions
// in LDS, block-wide reductions, no launches between iterations.
//
// Reductions.  Every block reduces its partial sums with wavefront shuffles
// in a fixed tree, writes them to a partials array, and the last block to
// finish (ticket counter) reduces the partials in fixed order and updates the
// device scalars (alpha, beta, r.z ...).  Deterministic, one launch per pass.
#include <hip/hip_runtime.h>
#include <cfloat>
#include <cmath>
#include <type_traits>

#include "hpcr_internal.h"
#include "hpcr_tile.h"
#include "hpcr_multi.h"

namespace hpcr {

thread_local Recorder* t_recorder = nullptr;

// ---- mechanical tile ops --------------------------------------------------
enum MechMode { M_APPLY = 0, M_RESID = 1, M_PCGA = 2, M_DIAG = 3 };

// FOLD: the grid is cubic (hx = hy = hz = h), so the gradient's 1/h and the
// transposed contraction's w_g/h fold into the moduli (times w_g/h^2, once
// per node plane on the finest level, once per element on coarse levels);
// with a power-of-two h this gives the same bits as the scaled form.
template <int MODE, bool FINE, bool FOLD = false>
struct MechOp {
  using T = double;
  using TR = double;
  static constexpr int NC = 3;
  static constexpr int NU = (MODE == M_DIAG) ? 0 : 3;  // vector values per node in LDS
  static constexpr int NK = FINE ? 3 : 0;              // K, G, d
  static constexpr int NV = (NU + NK) > 0 ? NU + NK : 1;
  static constexpr int NR = (MODE == M_PCGA) ? 6 + NK : NV;  // PCGA fetches z and p_old
  using Raw = double[NR];
  static constexpr int NSL = NSLOT, MIN_WAVES = HPCR_MECH_WAVES;  // LDS plane slots, occupancy target
  struct Acc {
    double a[3];
    __device__ Acc() { a[0] = a[1] = a[2] = 0.0; }
  };
  struct State {
    double beta;
  };
  // per-element data, staged through LDS with the node planes: the split
  // flag on the finest level, K_e and G_e on coarse levels
  using TE = std::conditional_t<FINE, int, double>;
  static constexpr int NE = FINE ? 1 : 2;
  MechCoef C;
  const double* x;      // APPLY/RESID: input vector; PCGA: z
  const double* invD;   // RESID
  const double* p_old;  // PCGA
  double* out;          // APPLY: y; RESID: r; PCGA: w; DIAG: invD
  double* out2;         // RESID: f; PCGA: p_new; DIAG: diag (or null)
  double* out3;         // RESID: z
  int flag;             // RESID: t0fine; PCGA: first
  double sx, sy, sz, tx_, ty_, tz_, wg, ihx2, ihy2, ihz2;
  double fs;            // FOLD: w_g / h^2
  Reduce R;
  size_t N;
  static_assert(!(FOLD && MODE == M_DIAG), "the diagonal keeps its own scaling");

  __device__ __forceinline__ State state() const {
    State s;
    s.beta = (MODE == M_PCGA && !flag) ? R.scal[S_BETA] : 0.0;
    return s;
  }

  // node values through buffer resources: a node outside the grid (or past
  // the chunk) reads zeros without a branch
  __device__ __forceinline__ void fetch(const LevelGeom& L, int gx, int gy, int gz, bool ok, const State& st,
                                        double (&r)[NR]) const {
#pragma unroll
    for (int k = 0; k < NR; k++) r[k] = 0.0;
    const unsigned ob = ok ? nid32(L, gx, gy, gz) * 8u : BOFF_NONE;
    const size_t nb = N * 8;
    if constexpr (MODE == M_APPLY || MODE == M_RESID || MODE == M_PCGA) {
#pragma unroll
      for (int c = 0; c < 3; c++) r[c] = bld(brsrc(x + c * N, nb), ob);
    }
    if constexpr (MODE == M_PCGA) {
      if (!flag) {
#pragma unroll
        for (int c = 0; c < 3; c++) r[3 + c] = bld(brsrc(p_old + c * N, nb), ob);
      }
    }
    if constexpr (FINE) {
      constexpr int o = NR - 3;
      r[o + 0] = bld(brsrc(C.K, nb), ob);
      r[o + 1] = bld(brsrc(C.G, nb), ob);
      r[o + 2] = bld(brsrc(C.d, nb), ob);
    }
  }

  __device__ __forceinline__ void efetch(const LevelGeom& L, int ex, int ey, int ez, bool ok, TE (&e)[NE]) const {
    const unsigned q = nid32(L, ex, ey, ez);
    if constexpr (FINE) {
      e[0] = bld_i8(brsrc(C.flag, N), ok ? q : BOFF_NONE);
    } else {
      const unsigned ob = ok ? 8u * q : BOFF_NONE;
      e[0] = bld(brsrc(C.Ke, N * 8), ob);
      e[1] = bld(brsrc(C.Ge, N * 8), ob);
    }
  }

  __device__ __forceinline__ void commit(const double (&r)[NR], const State& st, double (&v)[NV]) const {
    if constexpr (MODE == M_PCGA) {
#pragma unroll
      for (int c = 0; c < 3; c++) v[c] = r[c] + st.beta * r[3 + c];
#pragma unroll
      for (int k = 0; k < NK; k++) v[3 + k] = r[6 + k];
    } else {
#pragma unroll
      for (int k = 0; k < NV; k++) v[k] = r[k];
    }
    if constexpr (FINE && FOLD) {
      // K and G times w_g/h^2 (zero stays zero: the zero rule is unchanged)
      v[NU + 0] *= fs;
      v[NU + 1] *= fs;
    }
  }

  __device__ __forceinline__ void lame(const double (&sN)[NSL][NV][PLANE], int s0, int tx, int ty, const TE (&e)[NE],
                                       double (&lam)[2][2][2], double (&mu)[2][2][2], double& lam0,
                                       double& mu0) const {
    if constexpr (FINE) {
      double Kn[2][2][2], Gn[2][2][2], dn[2][2][2];
      gather8(sN, s0, tx, ty, NU + 0, Kn);
      gather8(sN, s0, tx, ty, NU + 1, Gn);
      gather8(sN, s0, tx, ty, NU + 2, dn);
      fine_lame(Kn, Gn, dn, e[0], C.small_k, lam, mu);
      lam0 = mu0 = 0.0;
    } else {
      lam0 = e[0] - (2.0 / 3.0) * e[1];
      mu0 = e[1];
      if constexpr (FOLD) {
        lam0 *= fs;
        mu0 *= fs;
      }
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
          for (int c = 0; c < 2; c++) {
            lam[a][b][c] = lam0;
            mu[a][b][c] = mu0;
          }
    }
  }

  __device__ __forceinline__ void element(const LevelGeom& L, const double (&sN)[NSL][NV][PLANE], int s0, int tx,
                                          int ty, int ex, int ey, int ez, const State& st, const TE (&e)[NE],
                                          double (&f)[3][2][2][2]) const {
    double lam[2][2][2], mu[2][2][2], lam0, mu0;
    lame(sN, s0, tx, ty, e, lam, mu, lam0, mu0);
    if constexpr (MODE == M_DIAG) {
      mech_element_diag(lam, mu, ihx2, ihy2, ihz2, wg, f);
    } else {
      double u[3][2][2][2];
#pragma unroll
      for (int c = 0; c < 3; c++) gather8(sN, s0, tx, ty, c, u[c]);
      if constexpr (FINE)
        mech_element<false, !FOLD>(u, lam, mu, 0.0, 0.0, sx, sy, sz, tx_, ty_, tz_, f);
      else
        mech_element<true, !FOLD>(u, lam, mu, lam0, mu0, sx, sy, sz, tx_, ty_, tz_, f);
    }
  }

  // finest-level operator passes: each element takes its bottom node plane's
  // interpolated moduli and in-plane gradient parts from the element below
  using Carry = std::conditional_t<FINE && MODE != M_DIAG && kDelta<double>, FineCarry<double>, void>;
  template <class C_, int PL>
  __device__ __forceinline__ void element_planes(const LevelGeom& L, const double (&sN)[NSL][NV][PL], int s0,
                                                 int tx, int ty, int ex, int ey, int ez, const State& st,
                                                 const TE (&e)[NE], double (&f)[3][2][2][2], C_& cr,
                                                 bool first) const {
    double Kn[2][2][2], Gn[2][2][2], dn[2][2][2];
    gather8(sN, s0, tx, ty, NU + 0, Kn);
    gather8(sN, s0, tx, ty, NU + 1, Gn);
    gather8(sN, s0, tx, ty, NU + 2, dn);
    double lam[2][2][2], mu[2][2][2];
    fine_lame_planes(Kn, Gn, dn, e[0], C.small_k, cr, first, lam, mu);
    double u[3][2][2][2];
#pragma unroll
    for (int c = 0; c < 3; c++) gather8(sN, s0, tx, ty, c, u[c]);
    mech_element_planes<!FOLD>(u, lam, mu, cr, first, sx, sy, sz, tx_, ty_, tz_, f);
  }

  static constexpr bool NEEDS_OWN = (MODE == M_PCGA);
  // RESID: f and 1/D of the owner's node, read one plane ahead
  struct NData {
    double f[MODE == M_RESID ? 3 : 1], iD[MODE == M_RESID ? 3 : 1];
  };
  __device__ __forceinline__ void nfetch(const LevelGeom& L, int i, int j, int k, NData& n) const {
    if constexpr (MODE == M_RESID) {
      const unsigned ob = nid32(L, i, j, k) * 8u;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        n.f[c] = ldo(out2 + c * N, ob);
        n.iD[c] = ldo(invD + c * N, ob);
      }
    }
  }
  // v: assembled values of node (i, j, k); own: the node's LDS values (p)
  __device__ __forceinline__ void epilogue(const LevelGeom& L, const double (&own)[3], int i, int j, int k, int dm,
                                           const double (&v)[3], const State& st, const NData& nd,
                                           Acc& acc) const {
    const unsigned ob = nid32(L, i, j, k) * 8u;
    if constexpr (MODE == M_APPLY) {
#pragma unroll
      for (int c = 0; c < 3; c++) sto(out + c * N, ob, v[c]);
    } else if constexpr (MODE == M_RESID) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        double fv = nd.f[c];
        const bool dir = (dm >> c) & 1;
        if (flag && dir) {
          fv = v[c];
          sto(out2 + c * N, ob, fv);
        }
        const double r = dir ? 0.0 : fv - v[c];
        const double z = r * nd.iD[c];
        sto(out + c * N, ob, r);
        sto(out3 + c * N, ob, z);
        acc.a[0] += r * z;
        acc.a[1] += r * r;
        acc.a[2] += fv * fv;
      }
    } else if constexpr (MODE == M_PCGA) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const double p = own[c];
        sto(out2 + c * N, ob, p);
        sto(out + c * N, ob, v[c]);
        acc.a[0] += p * v[c];
      }
    } else {  // M_DIAG
#pragma unroll
      for (int c = 0; c < 3; c++) {
        if (out2) sto(out2 + c * N, ob, v[c]);
        sto(out + c * N, ob, 1.0 / v[c]);
        acc.a[0] += (v[c] == 0.0) ? 1.0 : 0.0;
      }
    }
  }

  // split-output pass A (tile_body_split): node (i, j, k) of the tile's own
  // region, written to p_new and w as by epilogue()
  __device__ __forceinline__ void epilogue_split(const LevelGeom& L, const double (&p)[3], int i, int j, int k,
                                                 const double (&v)[3], Acc& acc) const {
    static_assert(MODE == M_PCGA, "split outputs are pass A's");
    const unsigned ob = nid32(L, i, j, k) * 8u;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      sto(out2 + c * N, ob, p[c]);
      sto(out + c * N, ob, v[c]);
      acc.a[0] += p[c] * v[c];
    }
  }
  // one component c of a side node's part, at byte offset so of the side
  // arrays; where no tile has the node in its own region (own_p: the last
  // node column or row of the grid) also p_new, and w = 0, at the node's byte
  // offset ob, so that pass B reads w at every node
  __device__ __forceinline__ void side1(const SplitW& S, unsigned so, double p, double v, bool own_p, unsigned ob,
                                        int c, Acc& acc) const {
    static_assert(MODE == M_PCGA, "split outputs are pass A's");
    sto(S.ws, so, v);
    acc.a[0] += p * v;
    if (own_p) {
      sto(out2 + c * N, ob, p);
      sto(out + c * N, ob, 0.0);
    }
  }

  __device__ __forceinline__ void finish(Acc& acc, const Blk& bk) const {
    if constexpr (MODE == M_APPLY) return;
    double* scal = R.scal;
    if constexpr (MODE == M_RESID) {
      const int t0 = flag;
      reduce_finalize<RED_SUM, 3>(bk, acc.a, R, 0.0, [=](double (&s)[3]) {
        scal[S_RZ] = s[0];
        scal[S_RR] = s[1];
        if (t0) scal[S_FF] = s[2];
      });
    } else if constexpr (MODE == M_PCGA) {
      double v1[1] = {acc.a[0]};
      reduce_finalize<RED_SUM, 1>(bk, v1, R, 0.0, [=](double (&s)[1]) {
        scal[S_PW] = s[0];
        scal[S_ALPHA] = scal[S_RZ] / s[0];
      });
    } else {
      double v1[1] = {acc.a[0]};
      reduce_finalize<RED_SUM, 1>(bk, v1, R, 0.0, [=](double (&s)[1]) { scal[S_ZERO] = s[0]; });
    }
  }
};

// ---- phase-field tile ops -------------------------------------------------
template <int MODE>
struct PfOp {
  using T = double;
  using TR = double;
  static constexpr int NC = 1;
  static constexpr int NU = (MODE == M_DIAG) ? 0 : 1;
  static constexpr int NV = NU + 1;  // + gc
  static constexpr int NR = (MODE == M_PCGA) ? 3 : NV;
  using Raw = double[NR];
  static constexpr int NSL = NSLOT, MIN_WAVES = 1;
  struct Acc {
    double a[3];
    __device__ Acc() { a[0] = a[1] = a[2] = 0.0; }
  };
  struct State {
    double beta;
  };
  // the element's strain history H, staged through LDS with the node planes
  using TE = double;
  static constexpr int NE = 1;
  __device__ __forceinline__ void efetch(const LevelGeom& L, int ex, int ey, int ez, bool ok, double (&e)[1]) const {
    e[0] = ok ? C.H[nidx(L, ex, ey, ez)] : 0.0;
  }
  PfCoef C;
  const double* x;      // APPLY/RESID: d; PCGA: z
  const double* invD;   // RESID
  const double* p_old;  // PCGA
  double* out;          // APPLY: y; RESID: R; PCGA: w; DIAG: invD
  double* out2;         // RESID: fd (read); PCGA: p_new; DIAG: diag or null
  double* out3;         // RESID: z
  int flag;             // PCGA: first
  double sx, sy, sz, wg, ihx2, ihy2, ihz2;
  Reduce R;

  __device__ __forceinline__ State state() const {
    State s;
    s.beta = (MODE == M_PCGA && !flag) ? R.scal[S_BETA] : 0.0;
    return s;
  }

  __device__ __forceinline__ void fetch(const LevelGeom& L, int gx, int gy, int gz, bool ok, const State& st,
                                        double (&r)[NR]) const {
#pragma unroll
    for (int k = 0; k < NR; k++) r[k] = 0.0;
    if (!ok) return;
    const size_t q = nidx(L, gx, gy, gz);
    if constexpr (MODE == M_APPLY || MODE == M_RESID) {
      r[0] = x[q];
    } else if constexpr (MODE == M_PCGA) {
      r[0] = x[q];
      if (!flag) r[1] = p_old[q];
    }
    r[NR - 1] = C.gc[q];
  }

  __device__ __forceinline__ void commit(const double (&r)[NR], const State& st, double (&v)[NV]) const {
    if constexpr (MODE == M_PCGA) {
      v[0] = r[0] + st.beta * r[1];
      v[1] = r[2];
    } else {
#pragma unroll
      for (int k = 0; k < NV; k++) v[k] = r[k];
    }
  }

  __device__ __forceinline__ void element(const LevelGeom& L, const double (&sN)[NSL][NV][PLANE], int s0, int tx,
                                          int ty, int ex, int ey, int ez, const State& st, const double (&e)[1],
                                          double (&f)[1][2][2][2]) const {
    double gcn[2][2][2];
    gather8(sN, s0, tx, ty, NU, gcn);
    double m[2][2][2], k[2][2][2];
    pf_coefs(gcn, e[0], C.lc, C.rlc, wg, m, k);
    if constexpr (MODE == M_DIAG) {
      pf_element_diag(m, k, ihx2, ihy2, ihz2, f[0]);
    } else {
      double v[2][2][2];
      gather8(sN, s0, tx, ty, 0, v);
      pf_element(v, m, k, sx, sy, sz, f[0]);
    }
  }

  static constexpr bool NEEDS_OWN = (MODE == M_PCGA);
  struct NData {
    double f, iD;
  };
  __device__ __forceinline__ void nfetch(const LevelGeom& L, int i, int j, int k, NData& n) const {
    if constexpr (MODE == M_RESID) {
      const size_t q = nidx(L, i, j, k);
      n.f = out2[q];
      n.iD = invD[q];
    }
  }
  __device__ __forceinline__ void epilogue(const LevelGeom& L, const double (&own)[1], int i, int j, int k, int dm,
                                           const double (&v)[1], const State& st, const NData& nd,
                                           Acc& acc) const {
    const size_t q = nidx(L, i, j, k);
    if constexpr (MODE == M_APPLY) {
      out[q] = v[0];
    } else if constexpr (MODE == M_RESID) {
      const double fv = nd.f;
      const double r = fv - v[0];
      const double z = r * nd.iD;
      out[q] = r;
      out3[q] = z;
      acc.a[0] += r * z;
      acc.a[1] += r * r;
      acc.a[2] += fv * fv;
    } else if constexpr (MODE == M_PCGA) {
      const double p = own[0];
      out2[q] = p;
      out[q] = v[0];
      acc.a[0] += p * v[0];
    } else {
      if (out2) out2[q] = v[0];
      out[q] = 1.0 / v[0];
      acc.a[0] += (v[0] == 0.0) ? 1.0 : 0.0;
    }
  }

  __device__ __forceinline__ void finish(Acc& acc, const Blk& bk) const {
    if constexpr (MODE == M_APPLY) return;
    double* scal = R.scal;
    if constexpr (MODE == M_RESID) {
      reduce_finalize<RED_SUM, 3>(bk, acc.a, R, 0.0, [=](double (&s)[3]) {
        scal[S_RZ] = s[0];
        scal[S_RR] = s[1];
        scal[S_FF] = s[2];
      });
    } else if constexpr (MODE == M_PCGA) {
      double v1[1] = {acc.a[0]};
      reduce_finalize<RED_SUM, 1>(bk, v1, R, 0.0, [=](double (&s)[1]) {
        scal[S_PW] = s[0];
        scal[S_ALPHA] = scal[S_RZ] / s[0];
      });
    } else {
      double v1[1] = {acc.a[0]};
      reduce_finalize<RED_SUM, 1>(bk, v1, R, 0.0, [=](double (&s)[1]) { scal[S_ZERO] = s[0]; });
    }
  }
};

static dim3 tile_grid(const LevelGeom& L) { return dim3((unsigned)tile_blocks(L.nx, L.ny, L.nz, L.zc)); }

template <int MODE, bool FINE, bool FOLD>
static MechOp<MODE, FINE, FOLD> mech_op(const LevelGeom& L, const MechCoef& C) {
  MechOp<MODE, FINE, FOLD> op{};
  op.C = C;
  op.N = (size_t)L.nx * L.ny * L.nz;
  op.sx = 1.0 / L.hx;
  op.sy = 1.0 / L.hy;
  op.sz = 1.0 / L.hz;
  op.wg = 0.125 * L.hx * L.hy * L.hz;
  op.tx_ = op.wg / L.hx;
  op.ty_ = op.wg / L.hy;
  op.tz_ = op.wg / L.hz;
  op.ihx2 = 1.0 / (L.hx * L.hx);
  op.ihy2 = 1.0 / (L.hy * L.hy);
  op.ihz2 = 1.0 / (L.hz * L.hz);
  op.fs = op.tx_ * op.sx;
  return op;
}

template <int MODE, bool FINE, bool FOLD>
static void run_mech_op(hipStream_t s, const LevelGeom& L, const MechCoef& C, const double* x, const double* invD,
                        const double* p_old, double* out, double* out2, double* out3, int flag, const Reduce& R,
                        const SplitW* S) {
  auto op = mech_op<MODE, FINE, FOLD>(L, C);
  op.x = x; op.invD = invD; op.p_old = p_old; op.out = out; op.out2 = out2; op.out3 = out3;
  op.flag = flag; op.R = R;
  using Op = MechOp<MODE, FINE, FOLD>;
  if constexpr (MODE == M_PCGA && FINE) {
    if (S) {
      const dim3 g((unsigned)split_blocks(L, *S));
      if (record<TileSplitK<Op>>(g, {L, op, *S})) return;
      tile_split_kernel<Op><<<g, dim3(TPB), 0, s>>>(L, op, *S);
      return;
    }
  }
  if (record<TileK<Op>>(tile_grid(L), {L, op})) return;
  tile_kernel<Op><<<tile_grid(L), dim3(TPB), 0, s>>>(L, op);
}

template <int MODE>
static void run_mech(hipStream_t s, const LevelGeom& L, const MechCoef& C, const double* x, const double* invD,
                     const double* p_old, double* out, double* out2, double* out3, int flag, const Reduce& R,
                     const SplitW* S = nullptr) {
  constexpr bool CAN_FOLD = (MODE != M_DIAG);
  const bool fold = CAN_FOLD && cubic_grid(L);
  if (C.fine) {
    if (fold)
      run_mech_op<MODE, true, CAN_FOLD>(s, L, C, x, invD, p_old, out, out2, out3, flag, R, S);
    else
      run_mech_op<MODE, true, false>(s, L, C, x, invD, p_old, out, out2, out3, flag, R, S);
  } else {
    if (fold)
      run_mech_op<MODE, false, CAN_FOLD>(s, L, C, x, invD, p_old, out, out2, out3, flag, R, S);
    else
      run_mech_op<MODE, false, false>(s, L, C, x, invD, p_old, out, out2, out3, flag, R, S);
  }
}

void launch_mech_apply(hipStream_t s, const LevelGeom& L, const MechCoef& C, const double* x, double* y) {
  Reduce R{};
  run_mech<M_APPLY>(s, L, C, x, nullptr, nullptr, y, nullptr, nullptr, 0, R);
}
void launch_mech_resid(hipStream_t s, const LevelGeom& L, const MechCoef& C, const double* u, double* f,
                       double* r, double* z, const double* invD, int t0fine, const Reduce& R) {
  run_mech<M_RESID>(s, L, C, u, invD, nullptr, r, f, z, t0fine, R);
}
void launch_mech_pcg_a(hipStream_t s, const LevelGeom& L, const MechCoef& C, const double* z,
                       const double* p_old, double* p_new, double* w, int first, const Reduce& R,
                       const SplitW* S) {
  run_mech<M_PCGA>(s, L, C, z, nullptr, p_old, w, p_new, nullptr, first, R, (S && S->ws && C.fine) ? S : nullptr);
}
void launch_mech_diag(hipStream_t s, const LevelGeom& L, const MechCoef& C, double* diag_or_null,
                      double* invD, const Reduce& R) {
  run_mech<M_DIAG>(s, L, C, nullptr, nullptr, nullptr, invD, diag_or_null, nullptr, 0, R);
}

template <int MODE>
static void run_pf(hipStream_t s, const LevelGeom& L, const PfCoef& C, const double* x, const double* invD,
                   const double* p_old, double* out, double* out2, double* out3, int flag, const Reduce& R) {
  PfOp<MODE> op{};
  op.C = C;
  op.x = x; op.invD = invD; op.p_old = p_old; op.out = out; op.out2 = out2; op.out3 = out3;
  op.flag = flag; op.R = R;
  op.sx = 1.0 / L.hx;
  op.sy = 1.0 / L.hy;
  op.sz = 1.0 / L.hz;
  op.wg = 0.125 * L.hx * L.hy * L.hz;
  op.ihx2 = 1.0 / (L.hx * L.hx);
  op.ihy2 = 1.0 / (L.hy * L.hy);
  op.ihz2 = 1.0 / (L.hz * L.hz);
  if (record<TileK<PfOp<MODE>>>(tile_grid(L), {L, op})) return;
  tile_kernel<PfOp<MODE>><<<tile_grid(L), dim3(TPB), 0, s>>>(L, op);
}

void launch_pf_apply(hipStream_t s, const LevelGeom& L, const PfCoef& C, const double* x, double* y) {
  Reduce R{};
  run_pf<M_APPLY>(s, L, C, x, nullptr, nullptr, y, nullptr, nullptr, 0, R);
}
void launch_pf_resid(hipStream_t s, const LevelGeom& L, const PfCoef& C, const double* d, const double* fd,
                     double* r, double* z, const double* invD, const Reduce& R) {
  run_pf<M_RESID>(s, L, C, d, invD, nullptr, r, const_cast<double*>(fd), z, 0, R);
}
void launch_pf_pcg_a(hipStream_t s, const LevelGeom& L, const PfCoef& C, const double* z, const double* p_old,
                     double* p_new, double* w, int first, const Reduce& R) {
  run_pf<M_PCGA>(s, L, C, z, nullptr, p_old, w, p_new, nullptr, first, R);
}
void launch_pf_diag(hipStream_t s, const LevelGeom& L, const PfCoef& C, double* diag_or_null, double* invD,
                    const Reduce& R) {
  run_pf<M_DIAG>(s, L, C, nullptr, nullptr, nullptr, invD, diag_or_null, nullptr, 0, R);
}

// ---------------------------------------------------------------------------
// Pointwise PCG pass B
// ---------------------------------------------------------------------------
constexpr int PW_TPB = 256;
static int pw_blocks(size_t n) {
  size_t b = (n + PW_TPB - 1) / PW_TPB;
  return (int)(b < 2048 ? (b == 0 ? 1 : b) : 2048);
}
// grid of the PCG pass-B kernels (at most HPCR_PASSB_MAX_BLOCKS blocks,
// grid-stride beyond)
#ifndef HPCR_PASSB_MAX_BLOCKS
#define HPCR_PASSB_MAX_BLOCKS 512
#endif
int passb_blocks(size_t n) {
  const size_t b = (n + PW_TPB - 1) / PW_TPB;
  const size_t cap = HPCR_PASSB_MAX_BLOCKS;
  return (int)(b == 0 ? 1 : (b < cap ? b : cap));
}

// Tuning macros of the pass-B kernels: HPCR_PASSB_NT (non-temporal loads and
// stores of the streamed vectors), HPCR_PASSB_UNROLL2 (two nodes per thread
// and iteration).
#ifdef HPCR_PASSB_NT
#define PB_LD(x) __builtin_nontemporal_load(&(x))
#define PB_ST(v, x) __builtin_nontemporal_store((v), &(x))
#else
#define PB_LD(x) (x)
#define PB_ST(v, x) ((x) = (v))
#endif

// Optional (tuning macro HPCR_PASSB_SPLIT): the pass-B kernels only write
// their block partials and a one-block kernel finishes the reduction, instead
// of the last-block ticket.
template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], const Reduce& R) {
  __shared__ double sh[NV * 16];
  block_red<RED_SUM, NV>(v, sh);
  if (threadIdx.x == 0)
    for (int k = 0; k < NV; k++) R.partials[(size_t)Blk::hw().lin() * NV + k] = v[k];
}
__global__ __launch_bounds__(256) void passb_finalize_kernel(const Reduce R, unsigned nb, int save_slot) {
  __shared__ double sh[2 * 16];
  double acc[2] = {0.0, 0.0};
  for (unsigned i = threadIdx.x; i < nb; i += blockDim.x) {
    acc[0] += R.partials[(size_t)i * 2];
    acc[1] += R.partials[(size_t)i * 2 + 1];
  }
  block_red<RED_SUM, 2>(acc, sh);
  if (threadIdx.x == 0) {
    double* scal = R.scal;
    scal[S_BETA] = acc[0] / scal[S_RZ];
    scal[S_RZ] = acc[0];
    scal[S_RR] = acc[1];
    if (save_slot >= 0) scal[S_AH + save_slot] = scal[S_ALPHA];
  }
}

// LDSF: the Dirichlet axis flags are staged in LDS (levels with
// nx + ny + nz <= DIRF_MAX), so a node's mask needs no dependent global load,
// which would also wait on the stores still in flight.
// NP: x updates of the CG group applied here (0: this iteration's is left
// out and its alpha kept in scal[S_AH + slot]).  A group's updates come in
// iteration order, each as the one fma of the undeferred form.
template <int NP>
struct GroupAlpha {
  double a[NP > 0 ? NP : 1];
  __device__ __forceinline__ GroupAlpha(const double* scal) {
#pragma unroll
    for (int m = 0; m + 1 < NP; m++) a[m] = scal[S_AH + m];
    a[NP > 0 ? NP - 1 : 0] = scal[S_ALPHA];
  }
  __device__ __forceinline__ double apply(const PGroup& P, size_t o, double x) const {
#pragma unroll
    for (int m = 0; m < NP; m++) x = fma(a[m], PB_LD(P.p[m][o]), x);
    return x;
  }
};

// w of node (i, j, k), component c, from the parts of a split-output pass A
// (SplitW), added in a fixed order: the part of the tile whose own region
// holds the node, of the tile below (top row), of the tile to its left (right
// column), of the tile below-left (corner).  A part is taken where that
// tile exists: none has the last node column x = 64 ntx or row y = 4 nty in
// its own region, nor a right column in row 4 nty, nor a top row in column
// 64 ntx.
__device__ __forceinline__ double split_w(const LevelGeom& L, const SplitW& S, const double* __restrict__ w, int c,
                                          size_t o, unsigned i, unsigned j, unsigned k) {
  const bool col = i < (unsigned)(STX * S.ntx), row = j < (unsigned)(STY * S.nty);
  const bool xe = i > 0 && i % STX == 0, ye = j > 0 && j % STY == 0;
  const unsigned kx = i / STX - 1, ky = j / STY - 1;
  // the own-region part is loaded unconditionally (pass A stores 0 at the
  // nodes past every own region); the right-column and corner parts, which
  // one lane in 64 needs, without a branch: an absent one from an offset past
  // the side arrays' resource, which reads 0 (and v + 0 = v)
  const auto rs = brsrc(S.ws, S.count() * 8);
  const unsigned ot = (3u * S.nr + c * S.nt + (k * (unsigned)S.nty + ky) * (unsigned)L.nx + i) * 8u;
  const unsigned orr = (c * S.nr + (k * (unsigned)L.ny + j) * (unsigned)S.ntx + kx) * 8u;
  const unsigned oc = (3u * (S.nr + S.nt) + c * S.nc + (k * (unsigned)S.nty + ky) * (unsigned)S.ntx + kx) * 8u;
  const double v = PB_LD(w[o]);
  double t = 0.0;
  if (ye && col) t = ldo(S.ws, ot);  // whole rows: a branch that waves mostly take or skip together
  const double r = bld(rs, (xe && row) ? orr : BOFF_NONE);
  const double q = bld(rs, (xe && ye) ? oc : BOFF_NONE);
  return ((v + t) + r) + q;
}

template <bool LDSF, int NP, bool WPARTS = false>
__device__ __forceinline__ void mech_pcg_b_body(const LevelGeom& L, double* __restrict__ u, double* __restrict__ r,
                                                double* __restrict__ z, const PGroup& P,
                                                const double* __restrict__ w, const double* __restrict__ invD,
                                                int slot, const Reduce& R, const SplitW& S, const Blk& bk) {
  __shared__ unsigned char sflags[LDSF ? DIRF_MAX : 1];
  const DirFlags F{sflags};
  if constexpr (LDSF) {
    F.stage(L);
    __syncthreads();
  }
  // node counts stay below 2^32 (a 1625^3 grid would not fit in HBM)
  const unsigned N = (unsigned)Nof(L);
  const unsigned plane = (unsigned)L.nx * (unsigned)L.ny;
  const double alpha = R.scal[S_ALPHA];
  const GroupAlpha<NP> ga(R.scal);
  double acc[2] = {0.0, 0.0};
  auto node = [&](unsigned q) {
    const unsigned k = q / plane;
    int dm = 0;
    unsigned i = 0, j = 0;
    if constexpr (LDSF || WPARTS) {
      const unsigned rem = q - k * plane;
      j = rem / (unsigned)L.nx;
      i = rem - j * (unsigned)L.nx;
    }
    if constexpr (LDSF) {
      dm = F.mask(L, (int)i, (int)j, (int)k);
    } else if (L.fz[k]) {
      const unsigned rem = q - k * plane;
      const unsigned jj = rem / (unsigned)L.nx, ii = rem - jj * (unsigned)L.nx;
      dm = dir_mask(L, (int)ii, (int)jj, (int)k);
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const size_t o = (size_t)c * N + q;
      if constexpr (NP > 0) PB_ST(ga.apply(P, o, PB_LD(u[o])), u[o]);
      double wv;
      if constexpr (WPARTS)
        wv = split_w(L, S, w, c, o, i, j, k);
      else
        wv = PB_LD(w[o]);
      double rv = PB_LD(r[o]) - alpha * wv;
      if ((dm >> c) & 1) rv = 0.0;
      PB_ST(rv, r[o]);
      const double zv = rv * PB_LD(invD[o]);
      PB_ST(zv, z[o]);
      acc[0] += rv * zv;
      acc[1] += rv * rv;
    }
  };
  const unsigned stride = bk.gx * PW_TPB;
#ifdef HPCR_PASSB_UNROLL2
  for (unsigned q = bk.x * PW_TPB + threadIdx.x; q < N; q += 2 * stride) {
    node(q);
    if (q + stride < N) node(q + stride);
  }
#else
  for (unsigned q = bk.x * PW_TPB + threadIdx.x; q < N; q += stride) node(q);
#endif
#ifdef HPCR_PASSB_SPLIT
  block_partials<2>(acc, R);
#else
  double* scal = R.scal;
  reduce_finalize<RED_SUM, 2>(bk, acc, R, 0.0, [=](double (&s)[2]) {
    scal[S_BETA] = s[0] / scal[S_RZ];
    scal[S_RZ] = s[0];
    scal[S_RR] = s[1];
    if constexpr (NP == 0) scal[S_AH + slot] = scal[S_ALPHA];
  });
#endif
}

template <bool LDSF, int NP, bool WPARTS>
__global__ __launch_bounds__(PW_TPB) void mech_pcg_b_kernel(const LevelGeom L, double* __restrict__ u,
                                                            double* __restrict__ r, double* __restrict__ z,
                                                            const PGroup P, const double* __restrict__ w,
                                                            const double* __restrict__ invD, int slot,
                                                            const Reduce R, const SplitW S) {
  mech_pcg_b_body<LDSF, NP, WPARTS>(L, u, r, z, P, w, invD, slot, R, S, Blk::hw());
}

struct PassBArgs {
  LevelGeom L;
  size_t N;
  double* x;  // u or d
  double* r;
  double* z;
  PGroup P;
  const double* w;
  const double* invD;
  int slot;
  Reduce R;
  SplitW S;   // mechanics: the parts of a split-output pass A (S.ws null: w whole)
};
template <bool LDSF, int NP, bool WPARTS>
struct MechPassBK {
  using Args = PassBArgs;
  static constexpr int NT = PW_TPB, MINW = 1;
  static dim3 block() { return dim3(PW_TPB); }
  __device__ __forceinline__ static void run(const Args& a, const Blk& bk) {
    mech_pcg_b_body<LDSF, NP, WPARTS>(a.L, a.x, a.r, a.z, a.P, a.w, a.invD, a.slot, a.R, a.S, bk);
  }
};

template <int NP>
__device__ __forceinline__ void pf_pcg_b_body(size_t N, double* __restrict__ d, double* __restrict__ r,
                                              double* __restrict__ z, const PGroup& P, const double* __restrict__ w,
                                              const double* __restrict__ invD, int slot, const Reduce& R,
                                              const Blk& bk) {
  const double alpha = R.scal[S_ALPHA];
  const GroupAlpha<NP> ga(R.scal);
  double acc[2] = {0.0, 0.0};
  for (size_t q = (size_t)bk.x * PW_TPB + threadIdx.x; q < N; q += (size_t)bk.gx * PW_TPB) {
    if constexpr (NP > 0) d[q] = ga.apply(P, q, d[q]);
    const double rv = r[q] - alpha * w[q];
    r[q] = rv;
    const double zv = rv * invD[q];
    z[q] = zv;
    acc[0] += rv * zv;
    acc[1] += rv * rv;
  }
  double* scal = R.scal;
  reduce_finalize<RED_SUM, 2>(bk, acc, R, 0.0, [=](double (&s)[2]) {
    scal[S_BETA] = s[0] / scal[S_RZ];
    scal[S_RZ] = s[0];
    scal[S_RR] = s[1];
    if constexpr (NP == 0) scal[S_AH + slot] = scal[S_ALPHA];
  });
}

template <int NP>
__global__ __launch_bounds__(PW_TPB) void pf_pcg_b_kernel(size_t N, double* __restrict__ d, double* __restrict__ r,
                                                          double* __restrict__ z, const PGroup P,
                                                          const double* __restrict__ w,
                                                          const double* __restrict__ invD, int slot,
                                                          const Reduce R) {
  pf_pcg_b_body<NP>(N, d, r, z, P, w, invD, slot, R, Blk::hw());
}
template <int NP>
struct PfPassBK {
  using Args = PassBArgs;
  static constexpr int NT = PW_TPB, MINW = 1;
  static dim3 block() { return dim3(PW_TPB); }
  __device__ __forceinline__ static void run(const Args& a, const Blk& bk) {
    pf_pcg_b_body<NP>(a.N, a.x, a.r, a.z, a.P, a.w, a.invD, a.slot, a.R, bk);
  }
};

template <bool LDSF, int NP, bool WPARTS>
static void mech_pcg_b_np(hipStream_t s, const LevelGeom& L, double* u, double* r, double* z, const PGroup& P,
                          const double* w, const double* invD, int slot, const Reduce& R, const SplitW& S) {
#ifndef HPCR_PASSB_SPLIT
  if (record<MechPassBK<LDSF, NP, WPARTS>>(dim3(passb_blocks(Nof(L))),
                                          PassBArgs{L, Nof(L), u, r, z, P, w, invD, slot, R, S}))
    return;
#endif
  hipLaunchKernelGGL((mech_pcg_b_kernel<LDSF, NP, WPARTS>), dim3(passb_blocks(Nof(L))), dim3(PW_TPB), 0, s, L, u, r,
                     z, P, w, invD, slot, R, S);
}
template <bool LDSF, bool WPARTS>
static void mech_pcg_b_ldsf(hipStream_t s, const LevelGeom& L, double* u, double* r, double* z, const PGroup& P,
                            const double* w, const double* invD, int np, int slot, const Reduce& R,
                            const SplitW& S) {
  static_assert(CG_GROUP == 4, "one instantiation per group position");
  switch (np) {
    case 0: mech_pcg_b_np<LDSF, 0, WPARTS>(s, L, u, r, z, P, w, invD, slot, R, S); break;
    case 1: mech_pcg_b_np<LDSF, 1, WPARTS>(s, L, u, r, z, P, w, invD, slot, R, S); break;
    case 2: mech_pcg_b_np<LDSF, 2, WPARTS>(s, L, u, r, z, P, w, invD, slot, R, S); break;
    case 3: mech_pcg_b_np<LDSF, 3, WPARTS>(s, L, u, r, z, P, w, invD, slot, R, S); break;
    default: mech_pcg_b_np<LDSF, 4, WPARTS>(s, L, u, r, z, P, w, invD, slot, R, S); break;
  }
}
void launch_mech_pcg_b(hipStream_t s, const LevelGeom& L, double* u, double* r, double* z, const PGroup& P,
                       const double* w, const double* invD, int np, int slot, const Reduce& R, const SplitW* S) {
  const bool parts = S && S->ws;
  const SplitW none{};
  const SplitW& SW = parts ? *S : none;
  if (L.nx + L.ny + L.nz <= DIRF_MAX) {
    if (parts)
      mech_pcg_b_ldsf<true, true>(s, L, u, r, z, P, w, invD, np, slot, R, SW);
    else
      mech_pcg_b_ldsf<true, false>(s, L, u, r, z, P, w, invD, np, slot, R, SW);
  } else {
    if (parts)
      mech_pcg_b_ldsf<false, true>(s, L, u, r, z, P, w, invD, np, slot, R, SW);
    else
      mech_pcg_b_ldsf<false, false>(s, L, u, r, z, P, w, invD, np, slot, R, SW);
  }
#ifdef HPCR_PASSB_SPLIT
  hipLaunchKernelGGL(passb_finalize_kernel, dim3(1), dim3(256), 0, s, R, (unsigned)passb_blocks(Nof(L)),
                     np == 0 ? slot : -1);
#endif
}
void launch_pf_pcg_b(hipStream_t s, const LevelGeom& L, double* d, double* r, double* z, const PGroup& P,
                     const double* w, const double* invD, int np, int slot, const Reduce& R) {
  const size_t N = Nof(L);
  const dim3 g(passb_blocks(N)), b(PW_TPB);
  const PassBArgs a{L, N, d, r, z, P, w, invD, slot, R};
  bool rec = false;
  switch (np) {
    case 0: rec = record<PfPassBK<0>>(g, a); break;
    case 1: rec = record<PfPassBK<1>>(g, a); break;
    case 2: rec = record<PfPassBK<2>>(g, a); break;
    case 3: rec = record<PfPassBK<3>>(g, a); break;
    default: rec = record<PfPassBK<4>>(g, a); break;
  }
  if (rec) return;
  switch (np) {
    case 0: hipLaunchKernelGGL(pf_pcg_b_kernel<0>, g, b, 0, s, N, d, r, z, P, w, invD, slot, R); break;
    case 1: hipLaunchKernelGGL(pf_pcg_b_kernel<1>, g, b, 0, s, N, d, r, z, P, w, invD, slot, R); break;
    case 2: hipLaunchKernelGGL(pf_pcg_b_kernel<2>, g, b, 0, s, N, d, r, z, P, w, invD, slot, R); break;
    case 3: hipLaunchKernelGGL(pf_pcg_b_kernel<3>, g, b, 0, s, N, d, r, z, P, w, invD, slot, R); break;
    default: hipLaunchKernelGGL(pf_pcg_b_kernel<4>, g, b, 0, s, N, d, r, z, P, w, invD, slot, R); break;
  }
}

// ---------------------------------------------------------------------------
// Dirichlet boundary conditions (boundary_conditions.cpp:28-111)
// ---------------------------------------------------------------------------
struct Planes {
  int k[4];
  int n;
};

__device__ __forceinline__ void dirichlet_bd_body(const LevelGeom& L, double* u, double Ut, const Planes& P,
                                                  const Blk& bk) {
  const size_t N = Nof(L);
  const int i = bk.x * blockDim.x + threadIdx.x;
  const int j = bk.y;
  if (i >= L.nx) return;
  // unrolled: the plane list is indexed with constants (no private copy)
#pragma unroll
  for (int t = 0; t < 4; t++) {
    if (t >= P.n) break;
    const int k = P.k[t];
    const int dm = dir_mask(L, i, j, k);
    const size_t q = nidx(L, i, j, k);
    if (dm & 1) u[q] = 0.0;
    if (dm & 2) u[N + q] = 0.0;
    if (dm & 4) u[2 * N + q] = (dm & 8) ? Ut : 0.0;
  }
}
__global__ void dirichlet_bd_kernel(const LevelGeom L, double* u, double Ut, Planes P) {
  dirichlet_bd_body(L, u, Ut, P, Blk::hw());
}
struct DirichletBdK {
  struct Args {
    LevelGeom L;
    double* u;
    double Ut;
    Planes P;
  };
  static constexpr int NT = 64, MINW = 1;
  static dim3 block() { return dim3(64); }
  __device__ __forceinline__ static void run(const Args& a, const Blk& bk) {
    dirichlet_bd_body(a.L, a.u, a.Ut, a.P, bk);
  }
};
__global__ void dirichlet_copy_kernel(const LevelGeom L, double* f, const double* Fi, Planes P) {
  const size_t N = Nof(L);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= L.nx) return;
  for (int t = 0; t < P.n; t++) {
    const int k = P.k[t];
    const int dm = dir_mask(L, i, j, k);
    const size_t q = nidx(L, i, j, k);
#pragma unroll
    for (int c = 0; c < 3; c++)
      if ((dm >> c) & 1) f[c * N + q] = Fi[c * N + q];
  }
}
static Planes planes_of(const int* kp, int n) {
  Planes P{};
  P.n = n > 4 ? 4 : n;
  for (int t = 0; t < P.n; t++) P.k[t] = kp[t];
  return P;
}
void launch_dirichlet_bd(hipStream_t s, const LevelGeom& L, double* u, double Ut, const int* kp, int n) {
  if (n <= 0) return;
  if (record<DirichletBdK>(dim3((L.nx + 63) / 64, L.ny), {L, u, Ut, planes_of(kp, n)})) return;
  hipLaunchKernelGGL(dirichlet_bd_kernel, dim3((L.nx + 63) / 64, L.ny), dim3(64), 0, s, L, u, Ut, planes_of(kp, n));
}
void launch_dirichlet_copy(hipStream_t s, const LevelGeom& L, double* f, const double* Fi, const int* kp, int n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(dirichlet_copy_kernel, dim3((L.nx + 63) / 64, L.ny), dim3(64), 0, s, L, f, Fi,
                     planes_of(kp, n));
}

// ---------------------------------------------------------------------------
// Whole smoother call of a small coarse level in one workgroup
// (pcg_u, mechanical_pcg.cpp:223-330; pcg_d, phase_field_pcg.cpp:133-214)
// ---------------------------------------------------------------------------
// Thread t owns nodes t and t + (block size) (u, f, r, z, 1/D and p in
// registers) and element t.  Per operator apply: p -> LDS, barrier, every
// element writes its 8 corner contributions to LDS, barrier, every node sums
// the contributions of its (up to) 8 elements in a fixed order.  Reductions
// are block-wide with a fixed shuffle tree.  The recurrence is that of the
// two-pass path: p = z + beta p, alpha = r.z / p.Ap, beta = r.z_new / r.z.

//
// A level of at most SMALL_WAVE_NODES nodes and 64 elements (the
// 4x4x4-element coarsest level) runs in one wavefront: its barriers and block
// sums then involve no other wave.
constexpr int SMALL_TPB = 256;
constexpr int SMALL_NPT = 2;  // nodes per thread
constexpr int SMALL_WAVE_NODES = 64 * SMALL_NPT;
static_assert(SMALL_NODES == SMALL_TPB * SMALL_NPT && SMALL_ELEMS == SMALL_TPB, "small-level layout");

// Block-wide sums, every thread receiving the total: wavefront partials to LDS,
// one barrier, then every thread adds the partials in wavefront order (the
// order of block_red, so the total is the same in every thread and equal to
// block_red's).  Consecutive calls alternate between the two halves of sh, so
// a half is written again only after the barrier of the call in between,
// which every thread reaches after its reads of that half.
template <int NV>
__device__ __forceinline__ void block_sum_all(double (&v)[NV], double* sh, int& half) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  double* h = sh + half * (4 * 16);
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = wave_red<RED_SUM>(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) h[k * 16 + wid] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double a = h[k * 16];
    for (int w = 1; w < nw; w++) a = a + h[k * 16 + w];
    v[k] = a;
  }
  half ^= 1;
}

template <class TIO>
struct SmallArgsT {
  LevelGeom L;
  // mechanics: element K_e, G_e; phase field: element H, nodal gc
  const double* Ke;
  const double* Ge;
  const double* H;
  const double* gc;
  double lc, rlc;      // lc and 1 / lc
  TIO* x;              // u (3 comps) or d
  const TIO* f;        // right-hand side
  const double* invD;  // inverse Jacobi diagonal
  TIO* r;              // final residual (optional)
  TIO* z;              // final r / D (optional)
  TIO* p;              // final search direction (optional)
  TIO* w;              // final A p (optional)
  double Ut;
  int NB;
  double* scal;
};
using SmallArgs = SmallArgsT<double>;

// TIO: type of the vectors in HBM; the iteration itself is FP64
template <int NC, class TIO, int TPB>
__device__ __forceinline__ void small_pcg_body(const SmallArgsT<TIO>& a) {
  const LevelGeom& L = a.L;
  __shared__ double sP[NC][SMALL_NODES];
  __shared__ double sE[NC * 8][SMALL_ELEMS];
  __shared__ double sh[2 * 4 * 16];
  int half = 0;
  const int tid = threadIdx.x;
  const int N = L.nx * L.ny * L.nz;
  const int ex_n = L.nx - 1, ey_n = L.ny - 1;
  const int E = ex_n * ey_n * (L.nz - 1);
  const double sx = 1.0 / L.hx, sy = 1.0 / L.hy, sz = 1.0 / L.hz;
  const double wg = 0.125 * L.hx * L.hy * L.hz;

  // element of this thread and its constant coefficients
  const bool has_e = tid < E;
  int ei = 0, ej = 0, ek = 0;
  if (has_e) {
    ei = tid % ex_n;
    ej = (tid / ex_n) % ey_n;
    ek = tid / (ex_n * ey_n);
  }
  double lam0 = 0.0, mu0 = 0.0;
  double pm[2][2][2], pk[2][2][2];
  if (has_e) {
    const size_t e = nidx(L, ei, ej, ek);
    if constexpr (NC == 3) {
      lam0 = a.Ke[e] - (2.0 / 3.0) * a.Ge[e];
      mu0 = a.Ge[e];
    } else {
      double gcn[2][2][2];
#pragma unroll
      for (int iz = 0; iz < 2; iz++)
#pragma unroll
        for (int iy = 0; iy < 2; iy++)
#pragma unroll
          for (int ix = 0; ix < 2; ix++) gcn[iz][iy][ix] = a.gc[nidx(L, ei + ix, ej + iy, ek + iz)];
      pf_coefs(gcn, a.H[e], a.lc, a.rlc, wg, pm, pk);
    }
  }

  // nodes of this thread
  int qn[SMALL_NPT], dm[SMALL_NPT];
  double xv[SMALL_NPT][NC], fv[SMALL_NPT][NC], iD[SMALL_NPT][NC];
#pragma unroll
  for (int s = 0; s < SMALL_NPT; s++) {
    const int q = tid + s * TPB;
    qn[s] = q < N ? q : -1;
    dm[s] = 0;
#pragma unroll
    for (int c = 0; c < NC; c++) xv[s][c] = fv[s][c] = iD[s][c] = 0.0;
    if (q < N) {
      if constexpr (NC == 3) {
        const int i = q % L.nx, j = (q / L.nx) % L.ny, k = q / (L.nx * L.ny);
        dm[s] = dir_mask(L, i, j, k);
      }
#pragma unroll
      for (int c = 0; c < NC; c++) {
        xv[s][c] = (double)a.x[(size_t)c * N + q];
        fv[s][c] = (double)a.f[(size_t)c * N + q];
        iD[s][c] = a.invD[(size_t)c * N + q];
      }
      // Dirichlet_BD (boundary_conditions.cpp:28-66)
      if constexpr (NC == 3) {
        if (dm[s] & 1) xv[s][0] = 0.0;
        if (dm[s] & 2) xv[s][1] = 0.0;
        if (dm[s] & 4) xv[s][2] = (dm[s] & 8) ? a.Ut : 0.0;
      }
    }
  }

  // y = A v for v in sP: element contributions, then node sums
  auto apply = [&](double (&y)[SMALL_NPT][NC]) {
    __syncthreads();
    if (has_e) {
      double f[NC][2][2][2];
      if constexpr (NC == 3) {
        double u[3][2][2][2];
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
          for (int iz = 0; iz < 2; iz++)
#pragma unroll
            for (int iy = 0; iy < 2; iy++)
#pragma unroll
              for (int ix = 0; ix < 2; ix++)
                u[c][iz][iy][ix] = sP[c][(ei + ix) + L.nx * ((ej + iy) + L.ny * (ek + iz))];
        double dl[2][2][2], dmu[2][2][2];
        mech_element<true>(u, dl, dmu, lam0, mu0, sx, sy, sz, wg / L.hx, wg / L.hy, wg / L.hz, f);
      } else {
        double v[2][2][2];
#pragma unroll
        for (int iz = 0; iz < 2; iz++)
#pragma unroll
          for (int iy = 0; iy < 2; iy++)
#pragma unroll
            for (int ix = 0; ix < 2; ix++) v[iz][iy][ix] = sP[0][(ei + ix) + L.nx * ((ej + iy) + L.ny * (ek + iz))];
        pf_element(v, pm, pk, sx, sy, sz, f[0]);
      }
#pragma unroll
      for (int c = 0; c < NC; c++)
#pragma unroll
        for (int iz = 0; iz < 2; iz++)
#pragma unroll
          for (int iy = 0; iy < 2; iy++)
#pragma unroll
            for (int ix = 0; ix < 2; ix++) sE[c * 8 + iz * 4 + iy * 2 + ix][tid] = f[c][iz][iy][ix];
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < SMALL_NPT; s++) {
#pragma unroll
      for (int c = 0; c < NC; c++) y[s][c] = 0.0;
      if (qn[s] < 0) continue;
      const int q = qn[s];
      const int i = q % L.nx, j = (q / L.nx) % L.ny, k = q / (L.nx * L.ny);
#pragma unroll
      for (int iz = 0; iz < 2; iz++)
#pragma unroll
        for (int iy = 0; iy < 2; iy++)
#pragma unroll
          for (int ix = 0; ix < 2; ix++) {
            const int e_i = i - ix, e_j = j - iy, e_k = k - iz;
            if (e_i < 0 || e_j < 0 || e_k < 0 || e_i >= ex_n || e_j >= ey_n || e_k >= L.nz - 1) continue;
            const int e = e_i + ex_n * (e_j + ey_n * e_k);
#pragma unroll
            for (int c = 0; c < NC; c++) y[s][c] += sE[c * 8 + iz * 4 + iy * 2 + ix][e];
          }
    }
  };

  // initial residual r = f - A x (masked), z = r / D
  double r[SMALL_NPT][NC], z[SMALL_NPT][NC], p[SMALL_NPT][NC], w[SMALL_NPT][NC];
#pragma unroll
  for (int s = 0; s < SMALL_NPT; s++)
    if (qn[s] >= 0)
#pragma unroll
      for (int c = 0; c < NC; c++) sP[c][qn[s]] = xv[s][c];
  apply(w);
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < SMALL_NPT; s++)
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const bool dir = (dm[s] >> c) & 1;
      r[s][c] = (qn[s] < 0 || dir) ? 0.0 : fv[s][c] - w[s][c];
      z[s][c] = r[s][c] * iD[s][c];
      p[s][c] = 0.0;
      acc[0] += r[s][c] * z[s][c];
      acc[1] += r[s][c] * r[s][c];
      acc[2] += fv[s][c] * fv[s][c];
    }
  block_sum_all<3>(acc, sh, half);
  // the phase-field smoother resets fd_norm on every call (phase_field_pcg.cpp:163)
  if (NC == 1 && tid == 0) a.scal[S_FF] = acc[2];
  double rz = acc[0], rr = acc[1], beta = 0.0, pw = 0.0, alpha = 0.0;

  for (int it = 0; it < a.NB; it++) {
    // pass A: p = z + beta p; w = A p; p.w
#pragma unroll
    for (int s = 0; s < SMALL_NPT; s++)
#pragma unroll
      for (int c = 0; c < NC; c++) {
        p[s][c] = (it == 0) ? z[s][c] : z[s][c] + beta * p[s][c];
        if (qn[s] >= 0) sP[c][qn[s]] = p[s][c];
      }
    apply(w);
    double pa[1] = {0.0};
#pragma unroll
    for (int s = 0; s < SMALL_NPT; s++)
#pragma unroll
      for (int c = 0; c < NC; c++) pa[0] += p[s][c] * w[s][c];
    block_sum_all<1>(pa, sh, half);
    pw = pa[0];
    alpha = rz / pw;
    // pass B: x += alpha p; r -= alpha w; mask; z = r / D; r.z, r.r
    double ab[2] = {0.0, 0.0};
#pragma unroll
    for (int s = 0; s < SMALL_NPT; s++)
#pragma unroll
      for (int c = 0; c < NC; c++) {
        xv[s][c] = xv[s][c] + alpha * p[s][c];
        double rv = r[s][c] - alpha * w[s][c];
        if ((dm[s] >> c) & 1) rv = 0.0;
        r[s][c] = rv;
        z[s][c] = rv * iD[s][c];
        ab[0] += rv * z[s][c];
        ab[1] += rv * rv;
      }
    block_sum_all<2>(ab, sh, half);
    beta = ab[0] / rz;
    rz = ab[0];
    rr = ab[1];
  }

#pragma unroll
  for (int s = 0; s < SMALL_NPT; s++) {
    if (qn[s] < 0) continue;
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const size_t o = (size_t)c * N + qn[s];
      a.x[o] = (TIO)xv[s][c];
      if (a.r) a.r[o] = (TIO)r[s][c];
      if (a.z) a.z[o] = (TIO)z[s][c];
      if (a.NB > 0 && a.p) {
        a.p[o] = (TIO)p[s][c];
        a.w[o] = (TIO)w[s][c];
      }
    }
  }
  if (tid == 0) {
    a.scal[S_RZ] = rz;
    a.scal[S_RR] = rr;
    if (a.NB > 0) {
      a.scal[S_BETA] = beta;
      a.scal[S_PW] = pw;
      a.scal[S_ALPHA] = alpha;
    }
  }
}

template <int NC, class TIO, int TPB>
__global__ __launch_bounds__(TPB) void small_pcg_kernel(const SmallArgsT<TIO> a) {
  small_pcg_body<NC, TIO, TPB>(a);
}
template <int NC, class TIO, int TPB>
struct SmallPcgK {
  using Args = SmallArgsT<TIO>;
  static constexpr int NT = TPB, MINW = 1;
  static dim3 block() { return dim3(TPB); }
  __device__ __forceinline__ static void run(const Args& a, const Blk&) { small_pcg_body<NC, TIO, TPB>(a); }
};
// one wavefront for the smallest levels, else SMALL_TPB threads
static bool small_wave(const LevelGeom& L) {
  return (long)L.nx * L.ny * L.nz <= SMALL_WAVE_NODES && (long)(L.nx - 1) * (L.ny - 1) * (L.nz - 1) <= 64;
}
template <int NC, class TIO>
static void launch_small(hipStream_t s, const LevelGeom& L, const SmallArgsT<TIO>& a) {
  if (small_wave(L)) {
    if (record<SmallPcgK<NC, TIO, 64>>(dim3(1), a)) return;
    small_pcg_kernel<NC, TIO, 64><<<dim3(1), dim3(64), 0, s>>>(a);
  } else {
    if (record<SmallPcgK<NC, TIO, SMALL_TPB>>(dim3(1), a)) return;
    small_pcg_kernel<NC, TIO, SMALL_TPB><<<dim3(1), dim3(SMALL_TPB), 0, s>>>(a);
  }
}

// coarsest solve of the FP32 preconditioner: zero start, homogeneous
// Dirichlet values, FP64 iteration on float right-hand side and result
void launch_small_pcg_u_f32(hipStream_t s, const LevelGeom& L, const double* Ke, const double* Ge, float* x,
                            const float* b, const double* invD, int NB, double* scal) {
  SmallArgsT<float> a{};
  a.L = L;
  a.Ke = Ke;
  a.Ge = Ge;
  a.x = x;
  a.f = b;
  a.invD = invD;
  a.Ut = 0.0;
  a.NB = NB;
  a.scal = scal;
  launch_small<3, float>(s, L, a);
}

bool small_level(const LevelGeom& L) {
  const long n = (long)L.nx * L.ny * L.nz;
  const long e = (long)(L.nx - 1) * (L.ny - 1) * (L.nz - 1);
  return n <= SMALL_NODES && e <= SMALL_ELEMS && L.nx >= 2 && L.ny >= 2 && L.nz >= 2;
}

void launch_small_pcg_u(hipStream_t s, const LevelGeom& L, const double* Ke, const double* Ge, double* u,
                        const double* f, const double* invD, double* r, double* z, double* p, double* w, double Ut,
                        int NB, double* scal) {
  SmallArgs a{};
  a.L = L;
  a.Ke = Ke;
  a.Ge = Ge;
  a.x = u;
  a.f = f;
  a.invD = invD;
  a.r = r;
  a.z = z;
  a.p = p;
  a.w = w;
  a.Ut = Ut;
  a.NB = NB;
  a.scal = scal;
  launch_small<3, double>(s, L, a);
}

void launch_small_pcg_d(hipStream_t s, const LevelGeom& L, const double* H, const double* gc, double lc, double* d,
                        const double* fd, const double* invD, double* r, double* z, double* p, double* w, int NB,
                        double* scal) {
  SmallArgs a{};
  a.L = L;
  a.H = H;
  a.gc = gc;
  a.lc = lc;
  a.rlc = 1.0 / lc;
  a.x = d;
  a.f = fd;
  a.invD = invD;
  a.r = r;
  a.z = z;
  a.p = p;
  a.w = w;
  a.NB = NB;
  a.scal = scal;
  launch_small<1, double>(s, L, a);
}

// ---------------------------------------------------------------------------
// image_MP on the device (image_input.cpp:120-142)
// ---------------------------------------------------------------------------
// gray level -> (K, G, gc) with the reference's integer-division
// interpolation between table rows, the same FP64 operations in the same
// order as the host restatement (no contraction), so results are identical
__global__ __launch_bounds__(256) void image_mp_kernel(size_t n, const unsigned char* __restrict__ img,
                                                       const MatTable* __restrict__ tab, double vs,
                                                       double* __restrict__ K, double* __restrict__ G,
                                                       double* __restrict__ gc, int* __restrict__ bad) {
#pragma clang fp contract(off)
  __shared__ int sg[MAT_MAX];
  __shared__ double sE[MAT_MAX], snu[MAT_MAX], sgc[MAT_MAX];
  const int nt = tab->n;
  for (int i = threadIdx.x; i < nt; i += blockDim.x) {
    sg[i] = tab->gray[i];
    sE[i] = tab->E[i];
    snu[i] = tab->nu[i];
    sgc[i] = tab->gc[i];
  }
  __syncthreads();
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
    const int x = img[q];
    int pos = -1;
    // findPosition (csv_reader.cpp:74-82)
    for (int i = 0; i < nt - 1; ++i)
      if (x >= sg[i] && x <= sg[i + 1]) {
        pos = i;
        break;
      }
    if (pos < 0 || sg[pos + 1] == sg[pos]) {
      bad[0] = 1;
      continue;
    }
    const int fac = (x - sg[pos]) / (sg[pos + 1] - sg[pos]);
    double E = sE[pos] + (double)fac * (sE[pos + 1] - sE[pos]);
    const double nu = snu[pos] + (double)fac * (snu[pos + 1] - snu[pos]);
    gc[q] = sgc[pos] + (double)fac * (sgc[pos + 1] - sgc[pos]);
    E = vs * E;
    K[q] = 1.0 / (3.0 * (1.0 - 2.0 * nu)) * E;
    G[q] = 1.0 / (2.0 * (1.0 + nu)) * E;
  }
}
void launch_image_mp(hipStream_t s, size_t n, const unsigned char* img, const MatTable* tab, double vs, double* K,
                     double* G, double* gc, int* bad) {
  hipLaunchKernelGGL(image_mp_kernel, dim3(pw_blocks(n)), dim3(256), 0, s, n, img, tab, vs, K, G, gc, bad);
}

// ---------------------------------------------------------------------------
// Transfers (multigrid_intergrid.cpp)
// ---------------------------------------------------------------------------
// solution_injection + Dirichlet_BD + copy to uold (multigrid_intergrid.cpp:25-121)
__device__ __forceinline__ void inject_body(const LevelGeom& Lc, const LevelGeom& Lf, int ncomp,
                                            const double* __restrict__ uf, double* __restrict__ uc,
                                            double* __restrict__ ucold, int apply_bc, double Ut, const Blk& bk) {
  const int i = bk.x * 64 + threadIdx.x;
  const int j = bk.y * 4 + threadIdx.y, k = bk.z;
  if (i >= Lc.nx || j >= Lc.ny) return;
  const size_t Nc = Nof(Lc), Nf = Nof(Lf);
  const size_t qc = nidx(Lc, i, j, k), qf = nidx(Lf, 2 * i, 2 * j, 2 * k);
  const int dm = apply_bc ? dir_mask(Lc, i, j, k) : 0;
  for (int c = 0; c < ncomp; c++) {
    double v = uf[c * Nf + qf];
    if ((dm >> c) & 1) v = (c == 2 && (dm & 8)) ? Ut : 0.0;
    uc[c * Nc + qc] = v;
    if (ucold) ucold[c * Nc + qc] = v;
  }
}
__global__ __launch_bounds__(256) void inject_kernel(const LevelGeom Lc, const LevelGeom Lf, int ncomp,
                                                     const double* __restrict__ uf, double* __restrict__ uc,
                                                     double* __restrict__ ucold, int apply_bc, double Ut) {
  inject_body(Lc, Lf, ncomp, uf, uc, ucold, apply_bc, Ut, Blk::hw());
}
struct InjectK {
  struct Args {
    LevelGeom Lc;
    LevelGeom Lf;
    int ncomp;
    const double* uf;
    double* uc;
    double* ucold;
    int apply_bc;
    double Ut;
  };
  static constexpr int NT = 256, MINW = 1;
  static dim3 block() { return dim3(64, 4); }
  __device__ __forceinline__ static void run(const Args& a, const Blk& bk) {
    inject_body(a.Lc, a.Lf, a.ncomp, a.uf, a.uc, a.ucold, a.apply_bc, a.Ut, bk);
  }
};
void launch_inject(hipStream_t s, const LevelGeom& Lc, const LevelGeom& Lf, int ncomp, const double* uf,
                   double* uc, double* ucold, int apply_bc, double Ut) {
  if (record<InjectK>(dim3((Lc.nx + 63) / 64, (Lc.ny + 3) / 4, Lc.nz), {Lc, Lf, ncomp, uf, uc, ucold, apply_bc, Ut}))
    return;
  hipLaunchKernelGGL(inject_kernel, dim3((Lc.nx + 63) / 64, (Lc.ny + 3) / 4, Lc.nz), dim3(64, 4), 0, s, Lc, Lf,
                     ncomp, uf, uc, ucold, apply_bc, Ut);
}

// 2-D blocks of 64 x 4 threads over (x, y), one z-plane per blockIdx.z
constexpr int XB = 64, YB = 4;

// coarsen_f_standard_routine (multigrid_intergrid.cpp:124-342): full weighting
// with out-of-range neighbours counted as 0, written in its separable form
// (1-D weights 1/2, 1, 1/2 per axis)
// one thread per node, x fastest: (i, j, k) of node q of a level
__device__ __forceinline__ void node_ijk(const LevelGeom& L, unsigned q, int& i, int& j, int& k) {
  const unsigned plane = (unsigned)L.nx * (unsigned)L.ny;
  const unsigned kk = q / plane, rem = q - kk * plane;
  const unsigned jj = rem / (unsigned)L.nx;
  i = (int)(rem - jj * (unsigned)L.nx);
  j = (int)jj;
  k = (int)kk;
}

__device__ __forceinline__ void restrict_body(const LevelGeom& Lc, const LevelGeom& Lf, int ncomp,
                                              const double* __restrict__ rf, const double* __restrict__ Fic,
                                              double* __restrict__ fc, const Blk& bk) {
  const unsigned qn = bk.x * 256 + threadIdx.x;
  if (qn >= (unsigned)Nof(Lc)) return;
  int i, j, k;
  node_ijk(Lc, qn, i, j, k);
  const size_t Nc = Nof(Lc), Nf = Nof(Lf);
  const Restrict27<double> R(Lf, i, j, k);
  const size_t qc = nidx(Lc, i, j, k);
  auto comps = [&](auto nc) {
    constexpr int NC = decltype(nc)::value;
    double acc[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) acc[c] = R(rf + c * Nf);
#pragma unroll
    for (int c = 0; c < NC; c++) fc[c * Nc + qc] = Fic[c * Nc + qc] + acc[c];
  };
  if (ncomp == 3)
    comps(std::integral_constant<int, 3>());
  else
    comps(std::integral_constant<int, 1>());
}
__global__ __launch_bounds__(256) void restrict_kernel(const LevelGeom Lc, const LevelGeom Lf, int ncomp,
                                                       const double* __restrict__ rf, const double* __restrict__ Fic,
                                                       double* __restrict__ fc) {
  restrict_body(Lc, Lf, ncomp, rf, Fic, fc, Blk::hw());
}
struct RestrictK {
  struct Args {
    LevelGeom Lc;
    LevelGeom Lf;
    int ncomp;
    const double* rf;
    const double* Fic;
    double* fc;
  };
  static constexpr int NT = 256, MINW = 1;
  static dim3 block() { return dim3(256); }
  __device__ __forceinline__ static void run(const Args& a, const Blk& bk) {
    restrict_body(a.Lc, a.Lf, a.ncomp, a.rf, a.Fic, a.fc, bk);
  }
};
void launch_restrict(hipStream_t s, const LevelGeom& Lc, const LevelGeom& Lf, int ncomp, const double* rf,
                     const double* Fic, double* fc) {
  if (record<RestrictK>(dim3((unsigned)((Nof(Lc) + 255) / 256)), {Lc, Lf, ncomp, rf, Fic, fc})) return;
  hipLaunchKernelGGL(restrict_kernel, dim3((unsigned)((Nof(Lc) + 255) / 256)), dim3(256), 0, s, Lc, Lf, ncomp, rf,
                     Fic, fc);
}

// e_c = u_c - u_c,old (the coarse-grid correction, multigrid_intergrid.cpp:484-518)
__device__ __forceinline__ void diff_body(size_t n, const double* __restrict__ a, const double* __restrict__ b,
                                          double* __restrict__ e, const Blk& bk) {
  for (size_t q = (size_t)bk.x * blockDim.x + threadIdx.x; q < n; q += (size_t)bk.gx * blockDim.x)
    e[q] = a[q] - b[q];
}
__global__ void diff_kernel(size_t n, const double* __restrict__ a, const double* __restrict__ b,
                            double* __restrict__ e) {
  diff_body(n, a, b, e, Blk::hw());
}
struct DiffK {
  struct Args {
    size_t n;
    const double* a;
    const double* b;
    double* e;
  };
  static constexpr int NT = 256, MINW = 1;
  static dim3 block() { return dim3(256); }
  __device__ __forceinline__ static void run(const Args& a, const Blk& bk) {
    diff_body(a.n, a.a, a.b, a.e, bk);
  }
};

// refine_standard_routine + Dirichlet_BD (multigrid_intergrid.cpp:416-549):
// trilinear interpolation of the correction e_c, one writer per fine node,
// two x-adjacent fine nodes per thread (ProlongPair)
__device__ __forceinline__ void prolong_body(const LevelGeom& Lf, const LevelGeom& Lc, int ncomp,
                                             double* __restrict__ uf, const double* __restrict__ e, int apply_bc,
                                             double Ut, const Blk& bk) {
  const unsigned q = bk.x * 256 + threadIdx.x;
  if (q >= (unsigned)prolong_pairs(Lf)) return;
  int ip, y, z;
  bool has1;
  pair_ijk(Lf, q, ip, has1, y, z);
  const size_t Nc = Nof(Lc), Nf = Nof(Lf);
  // separable trilinear weights (0.5 on each odd axis); identical to the
  // reference's averages of 2, 4 or 8 corrections up to rounding order
  const ProlongPair<double> Pr(Lc, ip, has1, y, z);
  const int x = 2 * ip;
  const unsigned of = nid32(Lf, x, y, z) * 8u;
  const int dm0 = apply_bc ? dir_mask(Lf, x, y, z) : 0;
  const int dm1 = (apply_bc && has1) ? dir_mask(Lf, x + 1, y, z) : 0;
  auto comps = [&](auto nc) {
    constexpr int NC = decltype(nc)::value;
    double c0[NC], c1[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) Pr(e + c * Nc, c0[c], c1[c]);
#pragma unroll
    for (int c = 0; c < NC; c++) {
      double v = ldo(uf + c * Nf, of) + c0[c];
      if ((dm0 >> c) & 1) v = (c == 2 && (dm0 & 8)) ? Ut : 0.0;
      sto(uf + c * Nf, of, v);
      if (has1) {
        double v1 = ldo(uf + c * Nf, of + 8u) + c1[c];
        if ((dm1 >> c) & 1) v1 = (c == 2 && (dm1 & 8)) ? Ut : 0.0;
        sto(uf + c * Nf, of + 8u, v1);
      }
    }
  };
  if (ncomp == 3)
    comps(std::integral_constant<int, 3>());
  else
    comps(std::integral_constant<int, 1>());
}
__global__ __launch_bounds__(256) void prolong_kernel(const LevelGeom Lf, const LevelGeom Lc, int ncomp,
                                                      double* __restrict__ uf, const double* __restrict__ e,
                                                      int apply_bc, double Ut) {
  prolong_body(Lf, Lc, ncomp, uf, e, apply_bc, Ut, Blk::hw());
}
struct ProlongK {
  struct Args {
    LevelGeom Lf;
    LevelGeom Lc;
    int ncomp;
    double* uf;
    const double* e;
    int apply_bc;
    double Ut;
  };
  static constexpr int NT = 256, MINW = 1;
  static dim3 block() { return dim3(256); }
  __device__ __forceinline__ static void run(const Args& a, const Blk& bk) {
    prolong_body(a.Lf, a.Lc, a.ncomp, a.uf, a.e, a.apply_bc, a.Ut, bk);
  }
};
void launch_prolong(hipStream_t s, const LevelGeom& Lf, const LevelGeom& Lc, int ncomp, double* uf,
                    const double* uc, const double* ucold, int apply_bc, double Ut, double* scratch) {
  const size_t n = (size_t)ncomp * Nof(Lc);
  if (record<DiffK>(dim3(pw_blocks(n)), {n, uc, ucold, scratch})) {
    record<ProlongK>(dim3((unsigned)((prolong_pairs(Lf) + 255) / 256)), {Lf, Lc, ncomp, uf, scratch, apply_bc, Ut});
    return;
  }
  hipLaunchKernelGGL(diff_kernel, dim3(pw_blocks(n)), dim3(256), 0, s, n, uc, ucold, scratch);
  hipLaunchKernelGGL(prolong_kernel, dim3((unsigned)((prolong_pairs(Lf) + 255) / 256)), dim3(256), 0, s, Lf, Lc, ncomp, uf,
                     scratch, apply_bc, Ut);
}

// ---------------------------------------------------------------------------
// Per-element Gauss-point kernels on the finest level
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load8(const LevelGeom& L, const double* f, int i, int j, int k,
                                      double (&v)[2][2][2]) {
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int c = 0; c < 2; c++) v[a][b][c] = f[nidx(L, i + c, j + b, k + a)];
}

struct Strain {
  double xx, yy, zz, xy, xz, yz;  // tensor components (xy = gamma_xy / 2)
};
__device__ __forceinline__ Strain strain_at(const double (&dx)[3][2][2], const double (&dy)[3][2][2],
                                            const double (&dz)[3][2][2], int gx, int gy, int gz) {
  Strain e;
  e.xx = dx[0][gz][gy];
  e.yy = dy[1][gz][gx];
  e.zz = dz[2][gy][gx];
  e.xy = 0.5 * (dx[1][gz][gy] + dy[0][gz][gx]);
  e.xz = 0.5 * (dx[2][gz][gy] + dz[0][gy][gx]);
  e.yz = 0.5 * (dy[2][gz][gx] + dz[1][gy][gx]);
  return e;
}
// epsilon_dev (mechanical_strain.cpp:47-61) and Phi_positive/negative
// (strain_history.cpp:134-159)
__device__ __forceinline__ double dev2(const Strain& e) {
  const double t = (1.0 / 3.0) * (e.xx + e.yy + e.zz);
  const double a = e.xx - t, b = e.yy - t, c = e.zz - t;
  return a * a + b * b + c * c + 2.0 * (e.xy * e.xy + e.xz * e.xz + e.yz * e.yz);
}
__device__ __forceinline__ double phi_pos(double K, double G, const Strain& e) {
  const double tr = e.xx + e.yy + e.zz;
  const double d2 = dev2(e);
  return (tr >= 0) ? 0.5 * K * tr * tr + G * d2 : G * d2;
}
__device__ __forceinline__ double phi_neg(double K, const Strain& e) {
  const double tr = e.xx + e.yy + e.zz;
  return (tr >= 0) ? 0.0 : 0.5 * K * tr * tr;
}

struct ElemArgs {
  const double* u;
  const double* K;
  const double* G;
  const double* d;
  const double* gc;
};

__device__ __forceinline__ void elem_grads(const LevelGeom& L, const double* u, int i, int j, int k,
                                           double (&dx)[3][2][2], double (&dy)[3][2][2], double (&dz)[3][2][2]) {
  const size_t N = Nof(L);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    double v[2][2][2];
    load8(L, u + c * N, i, j, k, v);
    grad8(v, 1.0 / L.hx, 1.0 / L.hy, 1.0 / L.hz, dx[c], dy[c], dz[c]);
  }
}

// MP_choice_element + MP_element (material_selection.cpp:19-45, 67-81)
__global__ void flag_means_kernel(const LevelGeom L, const double* __restrict__ u, const double* __restrict__ K,
                                  const double* __restrict__ G, const double* __restrict__ d,
                                  signed char* __restrict__ flag, double* __restrict__ Ke, double* __restrict__ Ge,
                                  double small_k, int compute_flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i > L.nx - 2) return;
  double Kn[2][2][2], Gn[2][2][2], dn[2][2][2], Kg[2][2][2], Gg[2][2][2], dg[2][2][2];
  load8(L, K, i, j, k, Kn);
  load8(L, G, i, j, k, Gn);
  gp_mp(Kn, Kg);
  gp_mp(Gn, Gg);
  const size_t e = nidx(L, i, j, k);
  int fl;
  if (compute_flag) {
    double dx[3][2][2], dy[3][2][2], dz[3][2][2];
    elem_grads(L, u, i, j, k, dx, dy, dz);
    double tr = 0.0, pp = 0.0, pn = 0.0;
    for (int gz = 0; gz < 2; gz++)
      for (int gy = 0; gy < 2; gy++)
        for (int gx = 0; gx < 2; gx++) {
          const Strain s = strain_at(dx, dy, dz, gx, gy, gz);
          tr += s.xx + s.yy + s.zz;
          pp += phi_pos(Kg[gz][gy][gx], Gg[gz][gy][gx], s);
          pn += phi_neg(Kg[gz][gy][gx], s);
        }
    fl = 0;
    if (tr >= 0.0) fl = 0;
    if (tr < 0.0) fl = 1;
    if (pp < pn) fl = 2;
    flag[e] = (signed char)fl;
  } else {
    fl = flag[e];
  }
  load8(L, d, i, j, k, dn);
  gp_mp(dn, dg);
  double sK = 0.0, sG = 0.0;
  for (int gz = 0; gz < 2; gz++)
    for (int gy = 0; gy < 2; gy++)
      for (int gx = 0; gx < 2; gx++) {
        const double g = degr(dg[gz][gy][gx], small_k);
        sK += (fl == 0) ? g * Kg[gz][gy][gx] : Kg[gz][gy][gx];
        sG += (fl == 2) ? Gg[gz][gy][gx] : g * Gg[gz][gy][gx];
      }
  Ke[e] = 0.125 * sK;
  Ge[e] = 0.125 * sG;
}
void launch_flag_means(hipStream_t s, const LevelGeom& L, const double* u, const double* K, const double* G,
                       const double* d, signed char* flag, double* Ke, double* Ge, double small_k,
                       int compute_flag) {
  hipLaunchKernelGGL(flag_means_kernel, dim3((L.nx - 1 + 63) / 64, L.ny - 1, L.nz - 1), dim3(64), 0, s, L, u, K, G,
                     d, flag, Ke, Ge, small_k, compute_flag);
}

// strain_history (strain_history.cpp:162-210)
__global__ void strain_history_kernel(const LevelGeom L, const double* __restrict__ u, const double* __restrict__ K,
                                      const double* __restrict__ G, const double* __restrict__ gc,
                                      double* __restrict__ H, double lc, double u_prop) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i > L.nx - 2) return;
  double Kn[2][2][2], Gn[2][2][2], cn[2][2][2], Kg[2][2][2], Gg[2][2][2], cg[2][2][2];
  load8(L, K, i, j, k, Kn);
  load8(L, G, i, j, k, Gn);
  load8(L, gc, i, j, k, cn);
  gp_mp(Kn, Kg);
  gp_mp(Gn, Gg);
  gp_mp(cn, cg);
  double dx[3][2][2], dy[3][2][2], dz[3][2][2];
  elem_grads(L, u, i, j, k, dx, dy, dz);
  double tmp = 0.0;
  for (int gz = 0; gz < 2; gz++)
    for (int gy = 0; gy < 2; gy++)
      for (int gx = 0; gx < 2; gx++) {
        Strain s = strain_at(dx, dy, dz, gx, gy, gz);
        s.xx *= u_prop; s.yy *= u_prop; s.zz *= u_prop;
        s.xy *= u_prop; s.xz *= u_prop; s.yz *= u_prop;
        tmp += phi_pos(Kg[gz][gy][gx], Gg[gz][gy][gx], s) - 0.5 * cg[gz][gy][gx] / lc;
      }
  const double Ht = 0.125 * tmp;
  const size_t e = nidx(L, i, j, k);
  if (H[e] < Ht) H[e] = Ht;
}
void launch_strain_history(hipStream_t s, const LevelGeom& L, const double* u, const double* K, const double* G,
                           const double* gc, double* H, double lc, double u_prop) {
  hipLaunchKernelGGL(strain_history_kernel, dim3((L.nx - 1 + 63) / 64, L.ny - 1, L.nz - 1), dim3(64), 0, s, L, u,
                     K, G, gc, H, lc, u_prop);
}

// U_e (strain_history.cpp:86-131): min over Gauss points of sqrt(phi_c / W)
__global__ __launch_bounds__(256) void elastic_limit_kernel(const LevelGeom L, const double* __restrict__ u,
                                                            const double* __restrict__ K,
                                                            const double* __restrict__ G,
                                                            const double* __restrict__ gc, double lc,
                                                            const Reduce R) {
  const size_t ne = (size_t)(L.nx - 1) * (L.ny - 1) * (L.nz - 1);
  double m[1] = {1e10};
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < ne; q += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(q % (L.nx - 1));
    const size_t r = q / (L.nx - 1);
    const int j = (int)(r % (L.ny - 1));
    const int k = (int)(r / (L.ny - 1));
    double Kn[2][2][2], Gn[2][2][2], cn[2][2][2], Kg[2][2][2], Gg[2][2][2], cg[2][2][2];
    load8(L, K, i, j, k, Kn);
    load8(L, G, i, j, k, Gn);
    load8(L, gc, i, j, k, cn);
    gp_mp(Kn, Kg);
    gp_mp(Gn, Gg);
    gp_mp(cn, cg);
    double dx[3][2][2], dy[3][2][2], dz[3][2][2];
    elem_grads(L, u, i, j, k, dx, dy, dz);
    for (int gz = 0; gz < 2; gz++)
      for (int gy = 0; gy < 2; gy++)
        for (int gx = 0; gx < 2; gx++) {
          const Strain e = strain_at(dx, dy, dz, gx, gy, gz);
          const double Kv = Kg[gz][gy][gx], Gv = Gg[gz][gy][gx];
          const double tr = e.xx + e.yy + e.zz;
          const double sxx = Kv * tr + 2.0 * Gv * (e.xx - 1.0 / 3.0 * tr);
          const double syy = Kv * tr + 2.0 * Gv * (e.yy - 1.0 / 3.0 * tr);
          const double szz = Kv * tr + 2.0 * Gv * (e.zz - 1.0 / 3.0 * tr);
          const double W = 0.5 * (sxx * e.xx + syy * e.yy + szz * e.zz +
                                  2.0 * (2.0 * Gv) * (e.xy * e.xy + e.xz * e.xz + e.yz * e.yz));
          const double phic = 0.5 * cg[gz][gy][gx] / lc;
          const double ut = sqrt(phic / W);
          if (ut < m[0]) m[0] = ut;
        }
  }
  double* scal = R.scal;
  reduce_finalize<RED_MIN, 1>(m, R, 1e10, [=](double (&s)[1]) { scal[S_RED] = s[0]; });
}
void launch_elastic_limit(hipStream_t s, const LevelGeom& L, const double* u, const double* K, const double* G,
                          const double* gc, double lc, const Reduce& R) {
  const size_t ne = (size_t)(L.nx - 1) * (L.ny - 1) * (L.nz - 1);
  hipLaunchKernelGGL(elastic_limit_kernel, dim3(pw_blocks(ne)), dim3(256), 0, s, L, u, K, G, gc, lc, R);
}

// ---------------------------------------------------------------------------
// Coarse coefficients (material_coarse_element.cpp, image_input.cpp)
// ---------------------------------------------------------------------------
__global__ void elem_vr_kernel(const LevelGeom Lc, const LevelGeom Lf, const double* __restrict__ Hf,
                               double* __restrict__ Hc, int voigt_only) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= Lc.nx) return;
  const size_t qc = nidx(Lc, i, j, k);
  const bool inside = i <= Lc.nx - 2 && j <= Lc.ny - 2 && k <= Lc.nz - 2;
  double V = 0.0, Rm = 0.0;
  if (inside) {
    const int x = 2 * i, y = 2 * j, z = 2 * k;
    const double h0 = Hf[nidx(Lf, x, y, z)], h1 = Hf[nidx(Lf, x + 1, y, z)];
    const double h2 = Hf[nidx(Lf, x, y + 1, z)], h3 = Hf[nidx(Lf, x + 1, y + 1, z)];
    const double h4 = Hf[nidx(Lf, x, y, z + 1)], h5 = Hf[nidx(Lf, x + 1, y, z + 1)];
    const double h6 = Hf[nidx(Lf, x, y + 1, z + 1)], h7 = Hf[nidx(Lf, x + 1, y + 1, z + 1)];
    V = 0.125 * (h0 + h1 + h2 + h3 + h4 + h5 + h6 + h7);
    if (!voigt_only)
      Rm = 8. / (1. / h0 + 1. / h1 + 1. / h2 + 1. / h3 + 1. / h4 + 1. / h5 + 1. / h6 + 1. / h7);
  }
  if (voigt_only) {
    if (inside) Hc[qc] = V;
  } else {
    Hc[qc] = 0.5 * (0.5 * (V + Rm) + sqrt(V * Rm));
  }
}
void launch_elem_vr(hipStream_t s, const LevelGeom& Lc, const LevelGeom& Lf, const double* Hf, double* Hc,
                    int voigt_only) {
  hipLaunchKernelGGL(elem_vr_kernel, dim3((Lc.nx + 63) / 64, Lc.ny, Lc.nz), dim3(64), 0, s, Lc, Lf, Hf, Hc,
                     voigt_only);
}

// voight / reuss / coarse_VR (image_input.cpp:144-273)
__global__ void nodal_vr_kernel(const LevelGeom Lc, const LevelGeom Lf, const double* __restrict__ M,
                                double* __restrict__ Mc) {
  const int ic = blockIdx.x * blockDim.x + threadIdx.x;
  const int jc = blockIdx.y, kc = blockIdx.z;
  if (ic >= Lc.nx) return;
  const int x = 2 * ic, y = 2 * jc, z = 2 * kc;
  const int xm = max(x - 1, 0), xp = min(x + 1, Lf.nx - 1);
  const int ym = max(y - 1, 0), yp = min(y + 1, Lf.ny - 1);
  const int zm = max(z - 1, 0), zp = min(z + 1, Lf.nz - 1);
  const int P[27][3] = {{xm, ym, zm}, {xm, ym, zp}, {xm, yp, zm}, {xm, yp, zp}, {xp, ym, zm}, {xp, ym, zp},
                        {xp, yp, zm}, {xp, yp, zp}, {xm, ym, z},  {xm, yp, z},  {xp, ym, z},  {xp, yp, z},
                        {xm, y, zm},  {xm, y, zp},  {xp, y, zm},  {xp, y, zp},  {x, ym, zm},  {x, ym, zp},
                        {x, yp, zm},  {x, yp, zp},  {x, y, zm},   {x, y, zp},   {x, ym, z},   {x, yp, z},
                        {xm, y, z},   {xp, y, z},   {x, y, z}};
  double sv = 0.0, sr = 0.0;
  for (int t = 0; t < 27; t++) {
    const double v = M[nidx(Lf, P[t][0], P[t][1], P[t][2])];
    sv += v;
    sr += 1. / v;
  }
  const double V = 1.0 / 27.0 * sv;
  const double R = 27.0 / sr;
  Mc[nidx(Lc, ic, jc, kc)] = 0.5 * (0.5 * (V + R) + sqrt(V * R));
}
void launch_nodal_vr(hipStream_t s, const LevelGeom& Lc, const LevelGeom& Lf, const double* Mf, double* Mc) {
  hipLaunchKernelGGL(nodal_vr_kernel, dim3((Lc.nx + 63) / 64, Lc.ny, Lc.nz), dim3(64), 0, s, Lc, Lf, Mf, Mc);
}

// fd_h (phase_field_right_hand_side.cpp:14-47) as a node gather
__global__ void fd_h_kernel(const LevelGeom L, const double* __restrict__ H, double* __restrict__ fd) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= L.nx) return;
  const double wg = 0.125 * L.hx * L.hy * L.hz;
  const double S = (GA + GB) * (GA + GB) * (GA + GB);  // sum_g N_a(g)
  double acc = 0.0;
  for (int dz = 1; dz >= 0; dz--)
    for (int dy = 1; dy >= 0; dy--)
      for (int dx = 1; dx >= 0; dx--) {
        const int ei = i - dx, ej = j - dy, ek = k - dz;
        if (ei < 0 || ej < 0 || ek < 0 || ei > L.nx - 2 || ej > L.ny - 2 || ek > L.nz - 2) continue;
        acc += wg * 2.0 * H[nidx(L, ei, ej, ek)] * S;
      }
  fd[nidx(L, i, j, k)] = acc;
}
void launch_fd_h(hipStream_t s, const LevelGeom& L, const double* H, double* fd) {
  hipLaunchKernelGGL(fd_h_kernel, dim3((L.nx + 63) / 64, L.ny, L.nz), dim3(64), 0, s, L, H, fd);
}

// externe_force_sum (mechanical_right_hnd.cpp:40-59)
__global__ __launch_bounds__(256) void top_sum_kernel(const LevelGeom L, const double* __restrict__ fz,
                                                      const Reduce R) {
  const size_t plane = (size_t)L.nx * L.ny;
  const size_t base = (size_t)(L.nz - 1) * plane;
  double a[1] = {0.0};
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < plane; q += (size_t)gridDim.x * blockDim.x)
    a[0] += fz[base + q];
  double* scal = R.scal;
  reduce_finalize<RED_SUM, 1>(a, R, 0.0, [=](double (&s)[1]) { scal[S_RED] = s[0]; });
}
void launch_top_sum(hipStream_t s, const LevelGeom& L, const double* fz, const Reduce& R) {
  const size_t plane = (size_t)L.nx * L.ny;
  hipLaunchKernelGGL(top_sum_kernel, dim3(pw_blocks(plane)), dim3(256), 0, s, L, fz, R);
}

// delta_d_max / d_max (phase_field_max_search.cpp:17-53): max(0, max(a - b))
__global__ __launch_bounds__(256) void max_diff_kernel(size_t n, const double* __restrict__ a,
                                                       const double* __restrict__ b, const Reduce R) {
  double m[1] = {0.0};
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
    const double v = b ? a[q] - b[q] : a[q];
    if (m[0] < v) m[0] = v;
  }
  double* scal = R.scal;
  reduce_finalize<RED_MAX, 1>(m, R, 0.0, [=](double (&s)[1]) { scal[S_RED] = s[0]; });
}
void launch_max_diff(hipStream_t s, size_t n, const double* a, const double* b_or_null, const Reduce& R) {
  hipLaunchKernelGGL(max_diff_kernel, dim3(pw_blocks(n)), dim3(256), 0, s, n, a, b_or_null, R);
}

// stiffness contrast of a nodal modulus field: max(a) into scal[S_RED] and
// max(1/a) over a > 0 into scal[S_SCRATCH] (a = 0 nodes are the zero-node
// rule's, not a phase)
__global__ __launch_bounds__(256) void contrast_kernel(size_t n, const double* __restrict__ a, const Reduce R) {
  double m[2] = {0.0, 0.0};
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
    const double v = a[q];
    if (m[0] < v) m[0] = v;
    if (v > 0.0 && m[1] < 1.0 / v) m[1] = 1.0 / v;
  }
  double* scal = R.scal;
  reduce_finalize<RED_MAX, 2>(m, R, 0.0, [=](double (&s)[2]) {
    scal[S_RED] = s[0];
    scal[S_SCRATCH] = s[1];
  });
}
void launch_contrast(hipStream_t s, size_t n, const double* a, const Reduce& R) {
  hipLaunchKernelGGL(contrast_kernel, dim3(pw_blocks(n)), dim3(256), 0, s, n, a, R);
}

// squared norm of ncomp*N values into scal[S_RR] (grid_norm, pcg_utilities.cpp:35-58)
__global__ __launch_bounds__(256) void norm2_kernel(size_t n, const double* __restrict__ a, const Reduce R) {
  double m[1] = {0.0};
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    m[0] += a[q] * a[q];
  double* scal = R.scal;
  reduce_finalize<RED_SUM, 1>(m, R, 0.0, [=](double (&s)[1]) { scal[S_RR] = s[0]; });
}
void launch_norm2(hipStream_t s, size_t n, const double* a, const Reduce& R) {
  hipLaunchKernelGGL(norm2_kernel, dim3(pw_blocks(n)), dim3(256), 0, s, n, a, R);
}

__global__ void flag_to_double_kernel(size_t n, const signed char* f, double* o) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    o[q] = (double)f[q];
}
__global__ void double_to_flag_kernel(size_t n, const double* in, signed char* f) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    f[q] = (signed char)in[q];
}
__device__ __forceinline__ void fill_body(size_t n, double* p, double v, const Blk& bk) {
  for (size_t q = (size_t)bk.x * blockDim.x + threadIdx.x; q < n; q += (size_t)bk.gx * blockDim.x)
    p[q] = v;
}
__global__ void fill_kernel(size_t n, double* p, double v) {
  fill_body(n, p, v, Blk::hw());
}
struct FillK {
  struct Args {
    size_t n;
    double* p;
    double v;
  };
  static constexpr int NT = 256, MINW = 1;
  static dim3 block() { return dim3(256); }
  __device__ __forceinline__ static void run(const Args& a, const Blk& bk) {
    fill_body(a.n, a.p, a.v, bk);
  }
};
void launch_flag_to_double(hipStream_t s, size_t n, const signed char* f, double* out) {
  hipLaunchKernelGGL(flag_to_double_kernel, dim3(pw_blocks(n)), dim3(256), 0, s, n, f, out);
}
void launch_double_to_flag(hipStream_t s, size_t n, const double* in, signed char* f) {
  hipLaunchKernelGGL(double_to_flag_kernel, dim3(pw_blocks(n)), dim3(256), 0, s, n, in, f);
}
void launch_fill(hipStream_t s, size_t n, double* p, double v) {
  if (record<FillK>(dim3(pw_blocks(n)), {n, p, v})) return;
  hipLaunchKernelGGL(fill_kernel, dim3(pw_blocks(n)), dim3(256), 0, s, n, p, v);
}

}  // namespace hpcr
