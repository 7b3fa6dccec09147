/*
 * hpcrack.h — C-ABI of the MI355X (gfx950) implementation of HPCrack's
 * matrix-free multigrid / Jacobi-PCG solves for degraded elasticity and the
 * AT2 phase field, plus the Gauss-point strain-history update between them.
 *
 * One context holds one problem: the whole level hierarchy of every field
 * lives in device memory (x-fastest, node-indexed, element fields stored in
 * node-sized grids whose last layer is unused, as the reference stores them).
 * Calls on one context come from one host thread.  Every function returns an
 * int status (HPCR_OK == 0); no exception crosses this boundary.
 *
 * Each entry point names the reference routine it replaces (paths relative to
 * the HPCrack source tree).  The reference has no FFI of its own: these are
 * its free C++ functions, re-exported with plain pointers and sizes.
 */
#ifndef HPCRACK_H
#define HPCRACK_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ----------------------------------------------------- */
#define HPCR_OK 0
#define HPCR_ERR_ARG (-1)           /* bad argument or shape                          */
#define HPCR_ERR_DEVICE (-2)        /* HIP runtime error / no device                  */
#define HPCR_ERR_MEMORY (-3)        /* device allocation failed                       */
#define HPCR_ERR_IO (-4)            /* file missing, short or malformed               */
#define HPCR_ERR_ZERO_DIAGONAL (-5) /* a Jacobi diagonal entry is 0: the reference
                                       throws "Division by zero in grid /= operation"
                                       (include/grid.h:215-217)                       */
#define HPCR_ERR_STATE (-6)         /* call order violated                            */

typedef struct hpcr_ctx hpcr_ctx;

/* MG cycle parameters, as MG (include/mg.h:16-23) */
typedef struct {
  int nu0;   /* relaxations on the coarsest level */
  int nu1;   /* pre-relaxations                   */
  int nu2;   /* post-relaxations                  */
  int gamma; /* 1: V-cycle, 2: W-cycle            */
  int ncy;   /* maximum cycles per solve          */
} hpcr_mg;

/* Level stack, as Stack/initialize (include/stack_and_level.h:15-39,
 * src/initialization.cpp:19-134) with one rank per axis. */
typedef struct {
  int maxlevel;      /* finest level index (src/main.cpp:52-68)            */
  int element_nb[3]; /* elements per axis on level 0 (config element_nb)   */
  double xa, xb, ya, yb, za, zb; /* computational domain                   */
} hpcr_geometry;

/* Physics and stopping constants that the reference hard-codes. */
typedef struct {
  double small_k;   /* residual stiffness k (src/crack_propagation.cpp:32: 1e-5)       */
  double lc;        /* phase-field length, already times h_fine (src/main.cpp:92-94)   */
  double tol_u;     /* squared relative residual of u (mechanical_displacement.cpp:88,102: 1e-6) */
  double tol_d_rel; /* squared relative residual of d (phase_field.cpp:43: 1e-12)      */
  double tol_d_abs; /* squared absolute residual of d (phase_field.cpp:30: 1e-12)      */
  hpcr_mg mg_u;     /* config ncy_u, nu0_u, ...                                       */
  hpcr_mg mg_d;     /* config ncy_d, nu0_d, ...                                       */
} hpcr_params;

#define HPCR_HIST_MAX 64
typedef struct {
  int cycles;                      /* MG cycles run by the solve                   */
  long long smoother_iters;        /* CG iterations, all levels                    */
  double r_norm_initial;           /* squared initial residual                     */
  double r_norm_final;             /* squared final residual                       */
  double history[HPCR_HIST_MAX];   /* squared residual after each cycle            */
} hpcr_stats;

/* Fields (x-fastest, N = nx*ny*nz doubles per level and component). */
enum hpcr_field {
  HPCR_U = 0,    /* displacement (3 components)                         */
  HPCR_FU = 1,   /* mechanical right-hand side (3)                      */
  HPCR_RU = 2,   /* mechanical residual (3)                             */
  HPCR_UOLD = 3, /* coarse solution before the coarse solve (3)         */
  HPCR_P = 4,    /* vector work field (3): operator input in tests      */
  HPCR_W = 5,    /* vector work field (3): operator output in tests;     
                    after a smoother on a finest level of 64 or more     
                    tiles per plane it holds the split pass A's own-tile 
                    parts, not the whole A p                             */
  HPCR_D = 6,    /* phase field                                         */
  HPCR_FD = 7,   /* phase-field right-hand side                         */
  HPCR_RD = 8,   /* phase-field residual                                */
  HPCR_DOLD = 9, /* coarse phase field before the coarse solve          */
  HPCR_H = 10,   /* strain history, element-indexed                     */
  HPCR_K = 11,   /* bulk modulus, nodal                                 */
  HPCR_G = 12,   /* shear modulus, nodal                                */
  HPCR_GC = 13,  /* fracture toughness, nodal                           */
  HPCR_KE = 14,  /* element bulk modulus (coarse operators)             */
  HPCR_GE = 15,  /* element shear modulus (coarse operators)            */
  HPCR_FLAG = 16,/* split flag 0/1/2, element-indexed, finest level     */
  HPCR_DPREV = 17,/* phase field of the previous load step, finest      */
  HPCR_PD = 18,  /* scalar work field: operator input in tests          */
  HPCR_WD = 19,  /* scalar work field: operator output in tests         */
  HPCR_FIELD_COUNT = 20
};

/* Timed kernels (hpcr_timing_query) */
enum hpcr_kernel {
  HPCR_K_MECH_PCG_A = 0, /* fused p-update + degraded-elasticity apply + p.Ap */
  HPCR_K_MECH_PCG_B = 1, /* fused u/r update + Dirichlet mask + r.z, r.r      */
  HPCR_K_MECH_RESID = 2, /* r = f - A u (smoother set-up)                      */
  HPCR_K_PF_PCG_A = 3,
  HPCR_K_PF_PCG_B = 4,
  HPCR_K_COUNT = 5
};

/* ---- context ---------------------------------------------------------- */
int hpcr_device_count(int* n);
int hpcr_ctx_create(const hpcr_geometry* geo, const hpcr_params* par, int device, hpcr_ctx** out);
int hpcr_ctx_destroy(hpcr_ctx* ctx);
const char* hpcr_last_error(const hpcr_ctx* ctx); /* NULL ctx: last creation error */
int hpcr_level_dims(const hpcr_ctx* ctx, int level, int dims[3]);
int hpcr_level_h(const hpcr_ctx* ctx, int level, double h[3]);
long long hpcr_device_bytes(const hpcr_ctx* ctx); /* device memory held */
/* Replay whole multigrid cycles of hpcr_u_entire / hpcr_phase_field as HIP
 * graphs (captured on first use per load value, one launch per cycle) instead
 * of launching their kernels one by one; results are identical.  Off by
 * default; kernel timing (hpcr_timing_enable) runs cycles eagerly. */
int hpcr_set_graphs(hpcr_ctx* ctx, int on);
/* The deterministic reductions (dot products, norms) order each block's
 * partial sums before its completion ticket by waiting for the stores, which
 * write through to the device-coherent L2 (the default, 0), or, with on = 1,
 * by a release/acquire fence pair -- the memory model's form, slower on this
 * GPU (it writes back the L2 once per block).  Results are identical. */
int hpcr_set_reduction_fence(hpcr_ctx* ctx, int on);

/* ---- fusion of several contexts' multigrid cycles ----------------------- */
/*
 * No counterpart in the reference (one MPI job solves one image).  For a
 * server holding many independent problems on one GPU: a context attached to
 * a fuser hands each multigrid cycle of hpcr_u_entire / hpcr_phase_field (and
 * each outer iteration of hpcr_u_entire_mixed) to the fuser's thread whenever
 * another attached context with the same level structure and size class is
 * inside a solve; otherwise it runs the cycle itself.  The fuser runs the
 * cycles submitted together (within window_us, cut short once no possible
 * partner is missing) as multi-problem launches: position i of the cycles of
 * such contexts is one launch per kernel type, every problem over its own
 * grid.  Results are bit-identical to unfused solves.  With graphs on (the
 * default) a batch is captured as a HIP graph when it recurs (same contexts,
 * same load) and replayed.  Calls on different contexts may come from
 * different threads; a context's own calls still come from one thread.
 */
typedef struct hpcr_fuser hpcr_fuser;
typedef struct {
  long long batches;          /* gathered batches                                */
  long long groups;           /* same-structure groups (one launch sequence each)  */
  long long cycles;           /* context cycles run through the fuser              */
  long long launches;         /* multi-problem launches issued                     */
  long long problem_launches; /* sum over those launches of their problems        */
  int max_problems;           /* most problems in one launch                       */
  long long graph_captures, graph_replays;
} hpcr_fuser_stats;
int hpcr_fuser_create(int device, int window_us, hpcr_fuser** out);
int hpcr_fuser_destroy(hpcr_fuser* fuser); /* HPCR_ERR_STATE while contexts are attached */
int hpcr_fuser_set_graphs(hpcr_fuser* fuser, int on);
int hpcr_fuser_get_stats(hpcr_fuser* fuser, hpcr_fuser_stats* out);
int hpcr_set_fuser(hpcr_ctx* ctx, hpcr_fuser* fuser); /* NULL detaches */
/* device bytes a context of this geometry holds (mixed == 1: plus the
 * buffers of the mixed-precision solve; mixed == 2: plus those of its FP64
 * preconditioner form), without allocating; for admission control of several
 * contexts on one device */
int hpcr_estimate_bytes(const hpcr_geometry* geo, int mixed, long long* bytes);
int hpcr_synchronize(hpcr_ctx* ctx);

/* Host <-> device field transfer: N doubles, x-fastest.  HPCR_FLAG takes
 * 0.0/1.0/2.0 values. */
int hpcr_set_field(hpcr_ctx* ctx, int field, int level, int comp, const double* host);
int hpcr_get_field(hpcr_ctx* ctx, int field, int level, int comp, double* host);
int hpcr_fill_field(hpcr_ctx* ctx, int field, int level, int comp, double value);

/* ---- set-up ------------------------------------------------------------- */
/* coarse_VR: 27-point Voigt/Reuss nodal coarsening of K, G or gc to all
 * coarse levels (src/image_input.cpp:144-273). */
int hpcr_coarse_vr(hpcr_ctx* ctx, int field);

/* ---- solve entry points (the drop-in boundary) ------------------------- */
/* u_entire (src/mechanical_displacement.cpp:25-115): split flag, element
 * moduli, coarse homogenisation, initial residual, up to ncy cycles to
 * sum(r^2)/f_norm <= tol_u, then the Dirichlet reactions into FU.  f_norm is
 * in/out and set during load step t == 0 only. */
int hpcr_u_entire(hpcr_ctx* ctx, double Ut, int t, double* f_norm, hpcr_stats* stats);
/* phase_field (src/phase_field.cpp:15-57) */
int hpcr_phase_field(hpcr_ctx* ctx, int t, double* fd_norm, hpcr_stats* stats);

/* Mixed-precision form of u_entire (same problem, boundary values, f_norm
 * convention, stopping test and reactions as hpcr_u_entire): FP64 flexible
 * CG on the finest level preconditioned by one FP32 multigrid V-cycle with
 * Chebyshev-Jacobi smoothing (degree cheb_degree before and after the coarse
 * correction, bounds [cheb_upper*lam/cheb_ratio, cheb_upper*lam] from
 * power_iters power iterations per level) and an FP64 coarsest solve of
 * coarse_iters (default nu0_u) Jacobi-PCG iterations.  Stops when sum(r^2)/f_norm <= tol with the residual
 * recomputed from u, or after max_iter iterations; the residual is also
 * recomputed every replace_every iterations (0: only at the end).  stats
 * counts outer iterations in `cycles`.  Allocates its FP32 buffers (about
 * 72 B per finest node plus 24 B FP64; element_mfma another 64 B per finest
 * element) on the first call. */
typedef struct hpcr_mixed_params {
  int max_iter;
  double tol;
  int cheb_degree; /* 1..8, or 0: chosen with cheb_ratio and cheb_upper from
                      the finest level's bulk-modulus contrast max K / min K
                      (K > 0): above 1e3 (near-void phases) degree 4, range
                      12, safety 1.3; otherwise degree 2, range 6, safety
                      1.1 (see hpcr_mixed_resolve) */
  int power_iters;
  double cheb_ratio;
  double cheb_upper;
  int replace_every;
  int precond_fp64; /* 0: the FP32 V-cycle (the mixed-precision form); 1: the
                       same V-cycle in FP64 -- the like-for-like comparison of
                       the two precisions at one tolerance */
  int element_mfma; /* FP32 V-cycle: 1 runs the finest level's element
                       contraction (B^T D B, src/mechanical_pcg.cpp:114-172) as
                       batched FP32 matrix-core products with FP32
                       accumulation; 0 (default) on the vector ALUs */
  int coarse_iters; /* Jacobi-PCG iterations of the preconditioner's coarsest
                       solve; 0 (default): the context's nu0_u */
} hpcr_mixed_params;
int hpcr_u_entire_mixed(hpcr_ctx* ctx, double Ut, int t, double* f_norm, const hpcr_mixed_params* mp,
                        hpcr_stats* stats);
/* The smoother setting hpcr_u_entire_mixed would use on the context's
 * current fields: with mp->cheb_degree == 0 fills cheb_degree, cheb_ratio and
 * cheb_upper from the modulus contrast (written to *contrast when not NULL);
 * otherwise leaves mp unchanged.  No counterpart in the reference. */
int hpcr_mixed_resolve(hpcr_ctx* ctx, hpcr_mixed_params* mp, double* contrast);

/* ---- between-solve updates --------------------------------------------- */
/* strain_history (src/strain_history.cpp:162-210), finest level */
int hpcr_strain_history(hpcr_ctx* ctx, double u_proportion);
/* Voigt_all_level (src/material_coarse_element.cpp:104-118), e.g. HPCR_H */
int hpcr_voigt_all_level(hpcr_ctx* ctx, int field);
/* VR_all_level (src/material_coarse_element.cpp:84-102), e.g. HPCR_KE */
int hpcr_vr_all_level(hpcr_ctx* ctx, int field);
/* fd_h (src/phase_field_right_hand_side.cpp:14-47) on one level */
int hpcr_fd_h(hpcr_ctx* ctx, int level);
/* U_e (src/strain_history.cpp:86-131) */
int hpcr_U_e(hpcr_ctx* ctx, double* Ue);
/* externe_force_sum (src/mechanical_right_hnd.cpp:40-59) on FU_z */
int hpcr_externe_force_sum(hpcr_ctx* ctx, double* sum);
/* delta_d_max / d_max (src/phase_field_max_search.cpp:17-53); d_old is HPCR_DPREV */
int hpcr_delta_d_max(hpcr_ctx* ctx, double* out);
int hpcr_d_max(hpcr_ctx* ctx, double* out);
int hpcr_save_d_old(hpcr_ctx* ctx); /* d_old = d (src/crack_propagation.cpp:93) */

/* ---- finer-grained entry points (operator and smoother level) ---------- */
/* Internal_Force (src/mechanical_pcg.cpp:174-191): dst = A(src), 3 comps */
int hpcr_internal_force(hpcr_ctx* ctx, int level, int src_field, int dst_field);
/* Internal_Force of the finest level in the mixed solve's FP32 preconditioner
 * (src/mechanical_pcg.cpp:174-191 at FP32): dst = -(A src), 0 on Dirichlet
 * DOFs, with the Gauss-point moduli from the current K, G, d and split flag;
 * mfma = 1 takes the matrix-core element contraction, 0 the vector one. */
int hpcr_internal_force_f32(hpcr_ctx* ctx, int src_field, int dst_field, int mfma);
/* PreM_u (src/mechanical_pcg.cpp:102-112): dst = diag(A) */
int hpcr_prem_u(hpcr_ctx* ctx, int level, int dst_field);
/* Fld (src/phase_field_pcg.cpp:93-103) and Precondition_M (:49-59) */
int hpcr_fld(hpcr_ctx* ctx, int level, int src_field, int dst_field);
int hpcr_precondition_m(hpcr_ctx* ctx, int level, int dst_field);
/* MP_choice (src/material_selection.cpp:55-65) on the finest level and
 * MP_all_element (:112-116) */
int hpcr_mp_choice(hpcr_ctx* ctx);
int hpcr_mp_all_element(hpcr_ctx* ctx);
/* set-up part of u_entire (src/mechanical_displacement.cpp:42-49) */
int hpcr_mech_setup(hpcr_ctx* ctx);
/* pcg_u (src/mechanical_pcg.cpp:223-330), pcg_d (src/phase_field_pcg.cpp:133-214) */
int hpcr_pcg_u(hpcr_ctx* ctx, int level, int nb_relax, double Ut, int t, double* f_norm);
int hpcr_pcg_d(hpcr_ctx* ctx, int level, int nb_relax, double* fd_norm);
/* cycle_u (src/mechanical_multigrid.cpp:18-54), cycle_d (src/phase_field_multigrid.cpp:18-44) */
int hpcr_cycle_u(hpcr_ctx* ctx, int level, double Ut, int t, double* f_norm);
int hpcr_cycle_d(hpcr_ctx* ctx, int level, double* fd_norm);
/* transfers (src/multigrid_intergrid.cpp:25-90, 124-342, 416-520) */
int hpcr_injection(hpcr_ctx* ctx, int coarse_level, int field, int comp);
int hpcr_restrict(hpcr_ctx* ctx, int fine_level, int rf_field, int fic_field, int dst_field, int comp);
int hpcr_prolong(hpcr_ctx* ctx, int fine_level, int uf_field, int uc_field, int ucold_field, int comp);
/* Dirichlet_BD (src/boundary_conditions.cpp:28-66) on a 3-component field */
int hpcr_dirichlet_bd(hpcr_ctx* ctx, int level, int field, double Ut);

/* ---- measurement ------------------------------------------------------- */
/* HIP-event timing of the kernels in enum hpcr_kernel launched on the
 * finest level, on the context's stream.  total_ms sums the recorded
 * launches since the last reset. */
int hpcr_timing_enable(hpcr_ctx* ctx, int on);
int hpcr_timing_reset(hpcr_ctx* ctx);
int hpcr_timing_query(hpcr_ctx* ctx, int kernel, int level, double* total_ms, long long* count);

#ifdef __cplusplus
}
#endif
#endif /* HPCRACK_H */
