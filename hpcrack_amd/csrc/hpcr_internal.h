// Internal declarations shared by the kernels (hpcr_kernels.hip) and the host
// orchestration (hpcr_solver.hip).  Not part of the C-ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace hpcr {

// Element-tile geometry of the operator kernels: 32x8 elements per layer,
// one element per thread, marching in z; a block owns the 31x7 interior node
// columns of its tile (neighbouring tiles overlap by one element).  32-wide
// rows keep the tile's global rows long (33 doubles) for cache-line use.
#ifndef HPCR_TILE_TX
#define HPCR_TILE_TX 32
#endif
constexpr int TX = HPCR_TILE_TX, TY = 256 / HPCR_TILE_TX;
constexpr int OWNX = TX - 1, OWNY = TY - 1;
constexpr int TPB = TX * TY;                  // 256 threads
constexpr int PX = TX + 1;                    // node row length in LDS
constexpr int PLANE = (TX + 1) * (TY + 1);    // 297 nodes per LDS plane (32x8 tiles)
static_assert(64 % TX == 0 && PLANE <= 2 * TX * TY, "tile rows must divide the wavefront");
constexpr int ZCHUNK_MAX = 64;                // node planes per block (fine levels)
// Split-output tiles of the finest level's fused PCG pass A: 64x4 elements
// per 256-thread block, a wave per 64-element row, every element computed
// once.  A tile writes all nodes of its 65x5 node region: the 64x4 at its
// elements' lower corners to w, its right column, top row and their corner
// (shared with the tiles to its right and above) to side arrays, which pass B
// adds in a fixed order.
constexpr int STX = 64, STY = 4;
constexpr int SPX = STX + 1;                  // node row length in LDS
constexpr int SPLANE = (STX + 1) * (STY + 1); // 325 nodes per LDS plane
static_assert(STX * STY == TPB && SPLANE <= 2 * TPB, "one element per thread, two plane nodes at most");
inline int split_tiles_x(int nx) { return (nx - 1 + STX - 1) / STX; }
inline int split_tiles_y(int ny) { return (ny - 1 + STY - 1) / STY; }
// The side arrays of a level with split pass A, 3 components each, one after
// the other in one allocation ws (element offsets 0, 3 nr, 3 (nr + nt)):
//   right columns x = 64 (kx + 1):    [c][z][y][kx],  nr = ntx * ny * nz
//   top rows y = 4 (ky + 1):          [c][z][ky][x],  nt = nx * nty * nz
//   their corners:                    [c][z][ky][kx], nc = ntx * nty * nz
// (only entries on the grid are written and read).  Byte offsets into ws fit
// in 32 bits as a level's own do: 3 (nr + nt + nc) < N on the levels that
// split (nr ~ N/64, nt ~ N/4), and a level holds at most MAX_LEVEL_NODES
// nodes.
struct SplitW {
  double* ws;
  int ntx, nty;
  int zc;  // node planes per block (the rounds model for the split tile count)
  unsigned nr, nt, nc;
  __host__ __device__ size_t count() const { return 3 * ((size_t)nr + nt + nc); }
};
// node planes held in LDS by a tile kernel: 3 lets one barrier per layer
// order the plane and exchange buffers, 2 needs two but less LDS (tuning
// macros: the FP64 passes, the FP32 preconditioner passes)
#ifndef HPCR_TILE_SLOTS
#define HPCR_TILE_SLOTS 3
#endif
#ifndef HPCR_MIXF_SLOTS
#define HPCR_MIXF_SLOTS 2
#endif
// minimum waves per SIMD the FP64 mechanical tile passes are compiled for
#ifndef HPCR_MECH_WAVES
#define HPCR_MECH_WAVES 1
#endif
// minimum waves per SIMD the FP32 preconditioner passes are compiled for
#ifndef HPCR_MIXF_WAVES
#define HPCR_MIXF_WAVES 1
#endif
constexpr int NSLOT = HPCR_TILE_SLOTS;
// A CG smoother call applies its x updates x += alpha_i p_i in groups: the
// pass B of iterations 0..CG_GROUP-2 of a group leaves the update out and
// keeps alpha_i, and the pass B that closes the group (or the call) applies
// the group's updates in iteration order.  The group's directions sit in
// CG_GROUP buffers, p_i in buffer i % CG_GROUP.
constexpr int CG_GROUP = 4;
struct PGroup {
  const double* p[CG_GROUP];
};
// levels small enough for the one-workgroup smoother (small_pcg_kernel)
constexpr int SMALL_NODES = 512, SMALL_ELEMS = 256;

// Device scalar slots (one set per context, reused by every level: the
// smoothers of different levels never overlap in time).
enum Scal {
  S_RZ = 0,    // r.z of the current CG iteration
  S_PW = 1,    // p.Ap
  S_ALPHA = 2,
  S_BETA = 3,
  S_RR = 4,    // r.r after the last update
  S_FF = 5,    // f.f (f_norm at t == 0)
  S_RED = 6,   // generic reduction result (min / max / sum)
  S_ZERO = 7,  // count of zero diagonal entries
  // flexible CG of the mixed-precision solve
  S_ZQ = 8,      // z.q_prev (preconditioned residual against the previous A p)
  S_PQ = 9,      // p.q
  S_PR = 10,     // p.r
  S_FALPHA = 11, // p.r / p.q
  S_LAM = 12,    // Rayleigh quotient of the power iteration
  S_AH = 13,     // CG_GROUP - 1 slots: alpha_i of the updates a smoother's group defers
  S_SCRATCH = 16,  // 8 slots written by the coarsest solve of the preconditioner
  S_APREV = 24,    // alpha of the flexible CG iteration whose x update is deferred
  S_COUNT = 25
};
static_assert(S_AH + CG_GROUP - 1 <= S_SCRATCH, "scalar slots of the deferred updates");

struct LevelGeom {
  int nx, ny, nz;
  int zc;  // node planes per block of the tile kernels (small levels use short chunks)
  double hx, hy, hz;
  const unsigned char* fx;  // Dirichlet axis flags: bit0 x == x_mid, bit1 x == x_b
  const unsigned char* fy;  // bit0 y == y_a,   bit1 y == y_mid
  const unsigned char* fz;  // bit0 z == z_a,   bit1 z == z_b
};

// most nodes of one level: the kernels address a node of one field
// component by a 32-bit byte offset (8 bytes per node), and a buffer
// resource's size (8 bytes per node) stays below the out-of-range offset
constexpr long long MAX_LEVEL_NODES = (1LL << 29) - 1;

// equal spacing in x, y and z: the mechanical tile passes fold the grid
// scale into the moduli
inline bool cubic_grid(const LevelGeom& L) { return L.hx == L.hy && L.hy == L.hz; }

struct Reduce {
  double* partials;     // >= blocks * 4 doubles
  unsigned int* counter; // zero between kernels
  double* scal;          // S_COUNT doubles
  int fence;             // order the partials before the ticket by a release fence (see reduce_finalize)
};

// Multi-problem launches.  While a recorder is set on the calling thread the
// launch functions record their kernels instead of issuing them; the fusion
// engine (hpcr_fuser) records one multigrid cycle of each of several contexts
// and issues position i of all their sequences as one grid, in which every
// problem keeps its own grid and argument record.
struct Rec;
using MultiFn = void (*)(hipStream_t, const Rec* const*, int);
constexpr int REC_ARG_BYTES = 448;
struct Rec {
  MultiFn fn;       // issues n records of one kernel type as one launch
  unsigned g[3];    // the problem's own grid
  alignas(16) unsigned char a[REC_ARG_BYTES];  // the kernel's argument record
};
struct Recorder {
  std::vector<Rec> recs;
};
// recorder of the calling thread (null: the launch functions launch)
extern thread_local Recorder* t_recorder;

// Mechanical coefficients of one level.
struct MechCoef {
  int fine;                   // 1: nodal K, G, d + split flag; 0: element K_e, G_e
  const double* K;
  const double* G;
  const double* d;
  const signed char* flag;
  const double* Ke;
  const double* Ge;
  double small_k;
};

// Coefficients of one level of the mixed-precision solve's preconditioner,
// whose arithmetic type P is float (the mixed-precision form) or double (the
// same V-cycle in FP64).  Finest level: the nodal K, G, d of type P (for P =
// double the FAS path's own fields) and the split flag of `m`, from which the
// passes form the Gauss-point moduli as the FAS passes do; with the
// matrix-core contraction, the precomputed FP32 Gauss-point moduli `lm`.
// Coarse levels: element moduli Ke, Ge of type P.
template <class P>
struct MixCoefT {
  int fine;
  int mfma;         // FP32 finest level: element contraction on the matrix cores
  const float* lm;  // FP32 finest level, matrix cores: (lambda, mu) at the 8 Gauss points, 16 floats per element
  const P* Kn;      // finest level: nodal K, G, d
  const P* Gn;
  const P* dn;
  const P* Ke;      // coarse levels: element moduli
  const P* Ge;
  MechCoef m;       // finest level: split flag, small_k
};

// Material table in device memory (layout of hpcr_material_table)
constexpr int MAT_MAX = 256;
struct MatTable {
  int n;
  int gray[MAT_MAX];
  double E[MAT_MAX], nu[MAT_MAX], gc[MAT_MAX];
};

// Phase-field coefficients of one level.
struct PfCoef {
  const double* H;   // element-indexed
  const double* gc;  // nodal
  double lc;
  double rlc;        // 1 / lc
};

// z-chunk of the tile kernels (node planes per block).
// Levels of at least ZC_MODEL_TILES tiles per plane (129^3 and up) are bound
// by throughput: the chunk count n minimises
//   ceil(blocks(n) / TILE_SLOTS) * (ceil(nz / n) + ZC_BLOCK_LAYERS),
// the rounds of resident blocks (two 256-thread blocks per CU on 256 CUs, the
// occupancy of the FP64 tile passes) times the element layers a block marches
// plus what its start and end cost in layers (the first element layer, the
// prologue's plane fetches, the reduction).  A last round of a few blocks
// otherwise leaves most CUs idle: at 257^3, 32-plane chunks give 2997 blocks
// (eight chunks of 32 planes and one of a single plane) in 6 rounds of 33
// layers, 43-plane chunks 1998 blocks in 4 rounds of 44.
// Smaller levels are bound by a block's latency: the longest power-of-two
// chunk up to ZC_SMALL_MAX that still gives >= HPCR_ZC_MIN_BLOCKS blocks, so
// that they are not serialised along z, else two planes, or one where that
// gives fewer than HPCR_ZC_ONE_BELOW blocks: those blocks then march over two
// element layers instead of three, which is what bounds such a pass while
// most compute units idle (17^3 and 9^3; on 33^3, 170 blocks, the extra
// element layers cost more than they save).
#ifndef HPCR_ZC_MIN_BLOCKS
#define HPCR_ZC_MIN_BLOCKS 1024
#endif
#ifndef HPCR_ZC_ONE_BELOW
#define HPCR_ZC_ONE_BELOW 128
#endif
constexpr int ZC_MODEL_TILES = 64, TILE_SLOTS = 512, ZC_BLOCK_LAYERS = 3, ZC_SMALL_MAX = 32;
inline long halo_tiles(int nx, int ny) { return (long)((nx + OWNX - 1) / OWNX) * ((ny + OWNY - 1) / OWNY); }
// the finest level's pass A runs on split-output tiles in the same regime
inline bool split_pass_a(int nx, int ny) { return halo_tiles(nx, ny) >= ZC_MODEL_TILES; }
// the chunk of the rounds model for xy tiles per plane
inline int model_zchunk(long xy, int nz) {
  long best = -1;
  int bz = 1;
  for (int n = (nz + ZCHUNK_MAX - 1) / ZCHUNK_MAX; n <= nz; n++) {
    const int zc = (nz + n - 1) / n;
    const long blocks = xy * ((nz + zc - 1) / zc);
    const long cost = ((blocks + TILE_SLOTS - 1) / TILE_SLOTS) * (zc + ZC_BLOCK_LAYERS);
    if (best < 0 || cost < best || (cost == best && zc < bz)) {
      best = cost;
      bz = zc;
    }
  }
  return bz;
}
inline int choose_zchunk(int nx, int ny, int nz) {
  const long xy = halo_tiles(nx, ny);
  if (xy >= ZC_MODEL_TILES) return model_zchunk(xy, nz);
  for (int zc = ZC_SMALL_MAX; zc > 2; zc /= 2)
    if (xy * ((nz + zc - 1) / zc) >= HPCR_ZC_MIN_BLOCKS) return zc;
  return xy * ((nz + 1) / 2) < HPCR_ZC_ONE_BELOW ? 1 : 2;
}
inline SplitW split_dims(int nx, int ny, int nz) {
  SplitW S{};
  S.ntx = split_tiles_x(nx);
  S.nty = split_tiles_y(ny);
  S.nr = (unsigned)S.ntx * (unsigned)ny * (unsigned)nz;
  S.nt = (unsigned)nx * (unsigned)S.nty * (unsigned)nz;
  S.nc = (unsigned)S.ntx * (unsigned)S.nty * (unsigned)nz;
  S.zc = model_zchunk((long)S.ntx * S.nty, nz);
  return S;
}
inline long tile_blocks(int nx, int ny, int nz, int zc) {
  return (long)((nx + OWNX - 1) / OWNX) * ((ny + OWNY - 1) / OWNY) * ((nz + zc - 1) / zc);
}
inline long split_blocks(const LevelGeom& L, const SplitW& S) {
  return (long)S.ntx * S.nty * ((L.nz + S.zc - 1) / S.zc);
}

// ---- launchers (hpcr_kernels.hip); all asynchronous on `s` -------------
// y = A x  (3 comps, stride N)
void launch_mech_apply(hipStream_t s, const LevelGeom& L, const MechCoef& C, const double* x, double* y);
// r = f - A u, masked, and z = r*invD; at t0fine also f := A u on Dirichlet
// DOFs; reductions r.z, r.r and f.f into scal.
void launch_mech_resid(hipStream_t s, const LevelGeom& L, const MechCoef& C, const double* u, double* f,
                       double* r, double* z, const double* invD, int t0fine, const Reduce& R);
// blocks of the pass-B launches over n nodes (sizes the reduction partials)
int passb_blocks(size_t n);
// p_new = z + beta p_old (first: p_new = z); w = A p_new; p.w -> alpha.
// S (finest level, with side arrays): split-output tiles, w in parts that
// pass B (given the same S) adds up
void launch_mech_pcg_a(hipStream_t s, const LevelGeom& L, const MechCoef& C, const double* z,
                       const double* p_old, double* p_new, double* w, int first, const Reduce& R,
                       const SplitW* S = nullptr);
// r -= alpha w; mask r; z = r*invD; r.z, r.r -> beta, rz, rr.  x updates:
// np == 0 keeps alpha in scal[S_AH + slot]; np >= 1 closes a group,
// u = fma(alpha_m, P.p[m], u) for m < np, alpha_m from scal[S_AH + m] and
// the last one the current alpha.
void launch_mech_pcg_b(hipStream_t s, const LevelGeom& L, double* u, double* r, double* z, const PGroup& P,
                       const double* w, const double* invD, int np, int slot, const Reduce& R,
                       const SplitW* S = nullptr);
// invD = 1 / diag(A); zero entries counted into scal[S_ZERO]; diag optionally kept
void launch_mech_diag(hipStream_t s, const LevelGeom& L, const MechCoef& C, double* diag_or_null,
                      double* invD, const Reduce& R);

void launch_pf_apply(hipStream_t s, const LevelGeom& L, const PfCoef& C, const double* x, double* y);
// R = fd - A d, z = R*invD; r.z, r.r, fd.fd
void launch_pf_resid(hipStream_t s, const LevelGeom& L, const PfCoef& C, const double* d, const double* fd,
                     double* r, double* z, const double* invD, const Reduce& R);
// the scalar forms of launch_mech_pcg_a / launch_mech_pcg_b
void launch_pf_pcg_a(hipStream_t s, const LevelGeom& L, const PfCoef& C, const double* z, const double* p_old,
                     double* p_new, double* w, int first, const Reduce& R);
void launch_pf_pcg_b(hipStream_t s, const LevelGeom& L, double* d, double* r, double* z, const PGroup& P,
                     const double* w, const double* invD, int np, int slot, const Reduce& R);
void launch_pf_diag(hipStream_t s, const LevelGeom& L, const PfCoef& C, double* diag_or_null, double* invD,
                    const Reduce& R);

// whole smoother call (Dirichlet values, residual, NB iterations) of a small
// coarse level in one workgroup; small_level() tells whether L fits
bool small_level(const LevelGeom& L);
void launch_small_pcg_u(hipStream_t s, const LevelGeom& L, const double* Ke, const double* Ge, double* u,
                        const double* f, const double* invD, double* r, double* z, double* p, double* w, double Ut,
                        int NB, double* scal);
void launch_small_pcg_d(hipStream_t s, const LevelGeom& L, const double* H, const double* gc, double lc, double* d,
                        const double* fd, const double* invD, double* r, double* z, double* p, double* w, int NB,
                        double* scal);

// ---- mixed-precision solve (hpcr_mixed.hip) ----
// The preconditioner's launchers are templated on its arithmetic type P
// (instantiated for float and double).
// Chebyshev step: r -= A s (masked), d = cd s + cr D^-1 r, x += d; with qprev,
// also x.qprev -> scal[S_ZQ]
template <class P>
void launch_mix_cheb(hipStream_t st, const LevelGeom& L, const MixCoefT<P>& C, const P* s, P* dout, P* r, P* x,
                     const P* invD, P cd, P cr, const double* qprev, const Reduce& R);
// r -= A s (masked)
template <class P>
void launch_mix_res(hipStream_t st, const LevelGeom& L, const MixCoefT<P>& C, const P* s, P* r);
// w = D^-1 A v (masked); Rayleigh quotient (v.Av)/(v.Dv) -> scal[S_LAM]
template <class P>
void launch_mix_power(hipStream_t st, const LevelGeom& L, const MixCoefT<P>& C, const P* v, P* w, const P* invD,
                      const Reduce& R);
template <class P>
void launch_power_init(hipStream_t st, const LevelGeom& L, P* v);
// flexible CG pass A (FP64): p = z + beta p_old, q = A p, p.q, p.r -> alpha.
// With u_lag (not first): also u += alpha_prev p_old, the update a lean
// pass B left out.
template <class P>
void launch_fcg_a(hipStream_t st, const LevelGeom& L, const MechCoef& C, const P* z, const double* p_old,
                  double* p_new, double* q, const double* r, int first, double* u_lag, const Reduce& R);
// the update a lean pass B left out, when no pass A follows: u += alpha_prev p
void launch_fcg_flush(hipStream_t st, const LevelGeom& L, double* u, const double* p, const Reduce& R);
// flexible CG pass B (FP64) + start of the next preconditioner application;
// lean: u and p are left alone and alpha is kept in scal[S_APREV]
template <class P>
void launch_fcg_b(hipStream_t st, const LevelGeom& L, double* u, double* r, const double* p, const double* q,
                  P* r32, P* d, P* x, const P* invD, P cr0, bool lean, const Reduce& R);
template <class P>
void launch_mix_start(hipStream_t st, const LevelGeom& L, const double* r64, P* r32, P* d, P* x, const P* invD,
                      P cr0);
template <class P>
void launch_mix_restrict(hipStream_t st, const LevelGeom& Lc, const LevelGeom& Lf, const P* rf, P* rc, P* dc, P* xc,
                         const P* invD, P cr0, int start);
template <class P>
void launch_mix_prolong(hipStream_t st, const LevelGeom& Lf, const LevelGeom& Lc, P* xf, P* cf, const P* ec);
// b = a converted to B (double -> float, float -> double, double -> double)
template <class A, class B>
void launch_convert(hipStream_t st, size_t n, const A* a, B* b);
// FP32 Gauss-point (lambda, mu) of the finest level's elements
void launch_mix_lame(hipStream_t st, const LevelGeom& L, const double* K, const double* G, const double* d,
                     const signed char* flag, double small_k, float* lm);
void launch_small_pcg_u_f32(hipStream_t s, const LevelGeom& L, const double* Ke, const double* Ge, float* x,
                            const float* b, const double* invD, int NB, double* scal);

// image_MP: gray levels -> K, G, gc; bad[0] := 1 for a level outside the table
void launch_image_mp(hipStream_t s, size_t n, const unsigned char* img, const MatTable* tab, double vs, double* K,
                     double* G, double* gc, int* bad);

// Dirichlet_BD on planes with z flags; `kplanes` lists the flagged k indices.
void launch_dirichlet_bd(hipStream_t s, const LevelGeom& L, double* u, double Ut, const int* kplanes_host,
                         int nplanes);
// f := Fi on Dirichlet DOFs
void launch_dirichlet_copy(hipStream_t s, const LevelGeom& L, double* f, const double* Fi,
                           const int* kplanes_host, int nplanes);
// injection u_c = u_f(2i), optional Dirichlet values, copy into uold
void launch_inject(hipStream_t s, const LevelGeom& Lc, const LevelGeom& Lf, int ncomp, const double* uf,
                   double* uc, double* ucold, int apply_bc, double Ut);
// f_c = Fic + full weighting of r_f
void launch_restrict(hipStream_t s, const LevelGeom& Lc, const LevelGeom& Lf, int ncomp, const double* rf,
                     const double* Fic, double* fc);
// u_f += P(u_c - u_c,old), optional Dirichlet values; `scratch` holds the
// coarse difference (ncomp * N_c doubles, may alias neither input)
void launch_prolong(hipStream_t s, const LevelGeom& Lf, const LevelGeom& Lc, int ncomp, double* uf,
                    const double* uc, const double* ucold, int apply_bc, double Ut, double* scratch);

// split flag + element means on the finest level
void launch_flag_means(hipStream_t s, const LevelGeom& L, const double* u, const double* K, const double* G,
                       const double* d, signed char* flag, double* Ke, double* Ge, double small_k,
                       int compute_flag);
// element VR (or Voigt only) coarsening from level f to level c (whole grid)
void launch_elem_vr(hipStream_t s, const LevelGeom& Lc, const LevelGeom& Lf, const double* Hf, double* Hc,
                    int voigt_only);
// nodal 27-point VR coarsening
void launch_nodal_vr(hipStream_t s, const LevelGeom& Lc, const LevelGeom& Lf, const double* Mf, double* Mc);
void launch_fd_h(hipStream_t s, const LevelGeom& L, const double* H, double* fd);
void launch_strain_history(hipStream_t s, const LevelGeom& L, const double* u, const double* K,
                           const double* G, const double* gc, double* H, double lc, double u_prop);
void launch_elastic_limit(hipStream_t s, const LevelGeom& L, const double* u, const double* K, const double* G,
                          const double* gc, double lc, const Reduce& R);
// reductions into scal[S_RED]
void launch_top_sum(hipStream_t s, const LevelGeom& L, const double* fz, const Reduce& R);
void launch_max_diff(hipStream_t s, size_t n, const double* a, const double* b_or_null, const Reduce& R);
void launch_contrast(hipStream_t s, size_t n, const double* a, const Reduce& R);

// squared norm into scal[S_RR]
void launch_norm2(hipStream_t s, size_t n, const double* a, const Reduce& R);
// flag conversions
void launch_flag_to_double(hipStream_t s, size_t n, const signed char* f, double* out);
void launch_double_to_flag(hipStream_t s, size_t n, const double* in, signed char* f);
void launch_fill(hipStream_t s, size_t n, double* p, double v);

}  // namespace hpcr
