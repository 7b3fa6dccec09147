// Host orchestration and C-ABI (include/hpcrack.h): one context per problem,
// the whole level hierarchy resident in device memory, every kernel on one
// HIP stream of default priority.  The control flow restates the reference's
// solve entry points (u_entire, phase_field), cycle drivers (cycle_u,
// cycle_d) and smoothers (pcg_u, pcg_d); the vector algebra of each smoother
// iteration runs in two fused passes whose scalars (alpha, beta) never leave
// the device.  The host synchronises once per multigrid cycle to read the
// residual norm for the stopping test.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hpcrack.h"
#include "../../include/hpcrack_io.h"
#include "hpcr_internal.h"

using namespace hpcr;

namespace {

thread_local std::string g_create_error;

struct Lev {
  LevelGeom g{};
  size_t N = 0;
  // mechanics (3 components each)
  double *u = nullptr, *fu = nullptr, *ru = nullptr, *uold = nullptr;
  double *p = nullptr, *p2 = nullptr, *w = nullptr, *invDu = nullptr, *zu = nullptr;
  double *p3 = nullptr, *p4 = nullptr;  // with p and p2: the directions of a CG group
  // phase field
  double *d = nullptr, *fd = nullptr, *rd = nullptr, *dold = nullptr;
  double *pd = nullptr, *pd2 = nullptr, *wd = nullptr, *invDd = nullptr, *zd = nullptr;
  double *pd3 = nullptr, *pd4 = nullptr;
  // coefficients
  double *H = nullptr, *K = nullptr, *G = nullptr, *gc = nullptr, *Ke = nullptr, *Ge = nullptr;
  signed char* flag = nullptr;
  double* dprev = nullptr;
  unsigned char *fx = nullptr, *fy = nullptr, *fz = nullptr;
  std::vector<int> kplanes;
  SplitW sw{};  // finest level: the side arrays of split-output pass A (ws null: not split)
};

// State of one level of the mixed-precision solve's preconditioner, in its
// arithmetic type P (float; or double for the FP64 V-cycle).  With P = double
// invD, Ke and Ge point at the FAS path's own FP64 buffers.
template <class P>
struct MixLev {
  P *x = nullptr, *r = nullptr, *dA = nullptr, *dB = nullptr, *invD = nullptr;
  float* lm = nullptr;  // finest level, matrix-core form: Gauss-point moduli (allocated on first use)
  P *Kn = nullptr, *Gn = nullptr, *dn = nullptr;  // finest level, FP32: nodal K, G, d
  P *Ke = nullptr, *Ge = nullptr;
  double lam = 0.0;
  P cr0 = P(0);
  std::vector<P> cd, cr;  // Chebyshev steps 1..m-1
};

struct TimingRec {
  int kernel, level;
  hipEvent_t a, b;
};

}  // namespace

struct hpcr_ctx {
  int device = 0;
  hpcr_geometry geo{};
  hpcr_params par{};
  int L = 0;  // maxlevel
  std::vector<Lev> lev;
  hipStream_t stream = nullptr;
  double* scal = nullptr;
  double* partials = nullptr;
  unsigned int* counter = nullptr;
  double* host_scal = nullptr;  // pinned
  size_t partial_cap = 0;
  long long bytes = 0;
  std::vector<void*> allocs;
  std::string err;
  bool timing = false;
  std::vector<TimingRec> trec;
  long long smoother_iters = 0;
  bool diag_u_valid = false, diag_d_valid = false;
  int last_rr_level = -1;  // level whose residual norm scal[S_RR] holds
  // HIP graphs of whole multigrid cycles (optional): captured once per
  // (load value, first-step flag) and replayed, one launch per cycle
  bool graphs = false;
  hipGraphExec_t g_u = nullptr, g_d = nullptr;
  double g_u_Ut = 0.0;
  int g_u_t0 = -1;
  long long g_u_iters = 0, g_d_iters = 0;  // smoother iterations per captured cycle
  // mixed-precision solve (allocated on first use): FP32 preconditioner, and
  // the FP64 one of the like-for-like comparison
  std::vector<MixLev<float>> mix;
  std::vector<MixLev<double>> mix64;
  double* q2 = nullptr;  // second FP64 q buffer of the flexible CG
  // fusion engine this context's multigrid cycles go through (optional)
  hpcr_fuser* fuser = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;  // ordering with the fused batches
  // counts the mixed-form set-ups: the Chebyshev coefficients, passed to the
  // kernels by value, change only there (a fused batch graph's key holds it)
  unsigned long long mix_epoch = 0;
  int mix_mfma = 0;  // the FP32 finest level's element contraction on the matrix cores
  int ticket_fence = 0;  // reductions order their partials by a release fence (hpcr_set_reduction_fence)
};

namespace {

#define HIPCHK(ctx, expr)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                      \
      return HPCR_ERR_DEVICE;                                                               \
    }                                                                                       \
  } while (0)

int set_err(hpcr_ctx* c, int code, const std::string& m) {
  c->err = m;
  return code;
}

// coordinate of index i (get_coordinate, boundary_conditions.cpp:13-17, one
// rank per axis); the product is rounded on its own, as in the reference.
double coord(double origin, double h, int i) {
  volatile double p = (double)i * h;
  return origin + p;
}

// Stream-ordered allocation on the context's stream: creating and destroying
// contexts does not synchronise the device, so contexts of other requests
// keep running meanwhile.
int dev_alloc(hpcr_ctx* c, void** p, size_t bytes) {
  hipError_t e = hipMallocAsync(p, bytes, c->stream);
  if (e != hipSuccess) {
    c->err = std::string("hipMallocAsync failed: ") + hipGetErrorString(e);
    *p = nullptr;
    return HPCR_ERR_MEMORY;
  }
  c->allocs.push_back(*p);
  c->bytes += (long long)bytes;
  e = hipMemsetAsync(*p, 0, bytes, c->stream);
  if (e != hipSuccess) {
    c->err = std::string("hipMemsetAsync failed: ") + hipGetErrorString(e);
    return HPCR_ERR_DEVICE;
  }
  return HPCR_OK;
}

template <class T>
int alloc_n(hpcr_ctx* c, T** p, size_t n) {
  return dev_alloc(c, (void**)p, n * sizeof(T));
}

Reduce red(hpcr_ctx* c) { return Reduce{c->partials, c->counter, c->scal, c->ticket_fence}; }

MechCoef mcoef(hpcr_ctx* c, int l) {
  Lev& v = c->lev[l];
  MechCoef m{};
  m.fine = (l == c->L) ? 1 : 0;
  m.K = v.K;
  m.G = v.G;
  m.d = v.d;
  m.flag = v.flag;
  m.Ke = v.Ke;
  m.Ge = v.Ge;
  m.small_k = c->par.small_k;
  return m;
}
PfCoef pcoef(hpcr_ctx* c, int l) { return PfCoef{c->lev[l].H, c->lev[l].gc, c->par.lc, 1.0 / c->par.lc}; }

int read_scal(hpcr_ctx* c) {
  if (t_recorder) return set_err(c, HPCR_ERR_STATE, "host read inside a recorded multigrid cycle");
  HIPCHK(c, hipMemcpyAsync(c->host_scal, c->scal, S_COUNT * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HPCR_OK;
}

// optional HIP-event bracket around one launch
struct TBracket {
  hpcr_ctx* c;
  int kernel, level;
  hipEvent_t a = nullptr, b = nullptr;
  TBracket(hpcr_ctx* c_, int k, int l) : c(c_), kernel(k), level(l) {
    if (!c->timing || l != c->L || t_recorder) return;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    (void)hipEventRecord(a, c->stream);
  }
  ~TBracket() {
    if (!a) return;
    (void)hipEventRecord(b, c->stream);
    c->trec.push_back(TimingRec{kernel, level, a, b});
  }
};

int check_launch(hpcr_ctx* c, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err(c, HPCR_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
  return HPCR_OK;
}

// ---- mechanics ------------------------------------------------------------
int mech_diag(hpcr_ctx* c, int l, double* diag_out) {
  launch_mech_diag(c->stream, c->lev[l].g, mcoef(c, l), diag_out, c->lev[l].invDu, red(c));
  return check_launch(c, "mech_diag");
}
int mech_diag_all(hpcr_ctx* c) {
  for (int l = 0; l <= c->L; l++) {
    int rc = mech_diag(c, l, nullptr);
    if (rc) return rc;
    if ((rc = read_scal(c))) return rc;
    if (c->host_scal[S_ZERO] != 0.0)
      return set_err(c, HPCR_ERR_ZERO_DIAGONAL, "Division by zero in grid /= operation (mechanical diagonal)");
  }
  c->diag_u_valid = true;
  return HPCR_OK;
}

// pcg_u (mechanical_pcg.cpp:223-330)
int pcg_u(hpcr_ctx* c, int l, int NB, double Ut, int t) {
  Lev& v = c->lev[l];
  int rc;
  if (!c->diag_u_valid) {
    if ((rc = mech_diag(c, l, nullptr))) return rc;
    if ((rc = read_scal(c))) return rc;
    if (c->host_scal[S_ZERO] != 0.0)
      return set_err(c, HPCR_ERR_ZERO_DIAGONAL, "Division by zero in grid /= operation (mechanical diagonal)");
  }
  if (l < c->L && small_level(v.g)) {
    launch_small_pcg_u(c->stream, v.g, v.Ke, v.Ge, v.u, v.fu, v.invDu, v.ru, v.zu, (NB & 1) ? v.p2 : v.p, v.w, Ut, NB,
                       c->scal);
    c->smoother_iters += NB;
    c->last_rr_level = l;
    return check_launch(c, "pcg_u (small level)");
  }
  launch_dirichlet_bd(c->stream, v.g, v.u, Ut, v.kplanes.data(), (int)v.kplanes.size());
  const int t0fine = (t == 0 && l == c->L) ? 1 : 0;
  {
    TBracket tb(c, HPCR_K_MECH_RESID, l);
    launch_mech_resid(c->stream, v.g, mcoef(c, l), v.u, v.fu, v.ru, v.zu, v.invDu, t0fine, red(c));
  }
  double* pb[CG_GROUP] = {v.p, v.p2, v.p3, v.p4};
  const PGroup P{{v.p, v.p2, v.p3, v.p4}};
  for (int i = 0; i < NB; i++) {
    const int m = i % CG_GROUP;
    {
      TBracket tb(c, HPCR_K_MECH_PCG_A, l);
      launch_mech_pcg_a(c->stream, v.g, mcoef(c, l), v.zu, pb[(m + CG_GROUP - 1) % CG_GROUP], pb[m], v.w, i == 0,
                        red(c), &v.sw);
    }
    {
      // u += alpha_i p_i waits for the pass B that closes the group
      const int np = (m == CG_GROUP - 1 || i == NB - 1) ? m + 1 : 0;
      TBracket tb(c, HPCR_K_MECH_PCG_B, l);
      launch_mech_pcg_b(c->stream, v.g, v.u, v.ru, v.zu, P, v.w, v.invDu, np, m, red(c), &v.sw);
    }
    c->smoother_iters++;
  }
  c->last_rr_level = l;
  return check_launch(c, "pcg_u");
}

// coarsen_u + coarsen_fu (multigrid_intergrid.cpp:106-121, 374-413)
int coarsen_u(hpcr_ctx* c, int lf, double Ut) {
  Lev& f = c->lev[lf];
  Lev& k = c->lev[lf - 1];
  launch_inject(c->stream, k.g, f.g, 3, f.u, k.u, k.uold, 1, Ut);
  launch_mech_apply(c->stream, k.g, mcoef(c, lf - 1), k.u, k.w);
  launch_restrict(c->stream, k.g, f.g, 3, f.ru, k.w, k.fu);
  return check_launch(c, "coarsen_u");
}

// cycle_u (mechanical_multigrid.cpp:18-54)
int cycle_u(hpcr_ctx* c, int l, double Ut, int t) {
  int rc;
  if (l == 0) return pcg_u(c, 0, c->par.mg_u.nu0, Ut, t);
  if ((rc = pcg_u(c, l, c->par.mg_u.nu1, Ut, t))) return rc;
  if ((rc = coarsen_u(c, l, Ut))) return rc;
  for (int j = 1; j <= c->par.mg_u.gamma; j++)
    if ((rc = cycle_u(c, l - 1, Ut, t))) return rc;
  Lev& f = c->lev[l];
  Lev& k = c->lev[l - 1];
  launch_prolong(c->stream, f.g, k.g, 3, f.u, k.u, k.uold, 1, Ut, k.w);
  return pcg_u(c, l, c->par.mg_u.nu2, Ut, t);
}

int mech_setup(hpcr_ctx* c) {
  Lev& F = c->lev[c->L];
  launch_flag_means(c->stream, F.g, F.u, F.K, F.G, F.d, F.flag, F.Ke, F.Ge, c->par.small_k, 1);
  for (int l = c->L; l > 0; l--) {
    launch_elem_vr(c->stream, c->lev[l - 1].g, c->lev[l].g, c->lev[l].Ke, c->lev[l - 1].Ke, 0);
    launch_elem_vr(c->stream, c->lev[l - 1].g, c->lev[l].g, c->lev[l].Ge, c->lev[l - 1].Ge, 0);
  }
  c->diag_u_valid = false;
  return check_launch(c, "mech_setup");
}

// Capture the kernel sequence of `body` on the context's stream into an
// executable graph.  The body must not synchronise or read back.
template <class F>
int capture_graph(hpcr_ctx* c, hipGraphExec_t* exec, F body) {
  if (*exec) {
    (void)hipGraphExecDestroy(*exec);
    *exec = nullptr;
  }
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = body();
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return set_err(c, HPCR_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    *exec = nullptr;
    return set_err(c, HPCR_ERR_DEVICE, std::string("graph instantiation: ") + hipGetErrorString(e));
  }
  return HPCR_OK;
}

int cycle_d(hpcr_ctx* c, int l);

}  // namespace

// ---- fusion engine ------------------------------------------------------------
// Multigrid cycles (and mixed-form outer iterations) of several contexts run as
// multi-problem launches.  A context attached to a fuser that has a partner --
// another attached context inside a solve loop with the same cycle signature
// -- hands each cycle to the fuser's thread and waits; without a partner it
// runs the cycle on its own stream as an unfused context does.  The thread
// gathers the cycles submitted within a window (cut short once no possible
// partner is missing), records each context's launch sequence (the same
// launch functions, with a recorder set), and issues position i of the
// sequences of a group of same-signature contexts as one launch per kernel
// type, on the first member's stream, ordered by events after and before the
// other members' streams.  Every problem keeps its own grid, argument record,
// reduction buffers and device scalars, so its results are bit-identical to a
// cycle of its own.  A batch is issued eagerly the first time its key (the
// members and the values they pass by value) is seen, and captured as a HIP
// graph and replayed when the key recurs.
struct FuseJob {
  hpcr_ctx* c = nullptr;
  std::string sig;             // contexts with equal signatures record aligned launch sequences
  std::string key;             // the recording's own parameters (with c: a batch graph's key)
  std::function<int()> body;   // records the context's launches (a cycle, an outer iteration)
  int rc = 0;
  long long iters = 0;
  bool done = false;
};

struct hpcr_fuser {
  int device = 0;
  int window_us = 200;
  bool graphs = true;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::vector<FuseJob*> pending;
  // attached contexts inside a solve's cycle loop, with the signature of the
  // last cycle each of them submitted (empty before the first)
  std::map<hpcr_ctx*, std::string> in_loop;
  int attached = 0;
  bool stop = false;
  std::thread th;
  struct Plan {
    std::vector<hpcr_ctx*> members;
    std::vector<long long> iters;  // smoother iterations of each member's recording
    hipGraphExec_t exec = nullptr;
    long long launches = 0, problem_launches = 0;
    int max_problems = 0;
    unsigned long long used = 0;
  };
  std::mutex plan_mu;  // plans: the fuser's thread and detaching contexts
  std::map<std::string, Plan> plans;
  std::map<std::string, int> seen;  // batch keys issued eagerly once (captured when they come back)
  unsigned long long tick = 0;
  hpcr_fuser_stats st{};
};

namespace {

constexpr size_t FUSE_PLAN_MAX = 64;  // cached batch graphs per fuser
constexpr int FUSE_GROUP_MAX = 8;     // problems per launch (MULTI_MAX of the kernels)

// Contexts with equal signatures record cycles whose positions correspond:
// same kind, level count, cycle parameters, and one-workgroup coarse levels.
// (kind 2: outer iteration of the mixed-precision solve, FP32 V-cycle with
// the mechanical cycle parameters)
std::string cycle_signature(hpcr_ctx* c, int kind) {
  const hpcr_mg& m = kind == 1 ? c->par.mg_d : c->par.mg_u;
  std::string s = std::to_string(kind) + ":" + std::to_string(c->L) + ":" + std::to_string(m.nu0) + "," +
                  std::to_string(m.nu1) + "," + std::to_string(m.nu2) + "," + std::to_string(m.gamma) + ":";
  for (int l = 0; l < c->L; l++) s += small_level(c->lev[l].g) ? 'S' : 'T';
  // and finest levels within a factor of two in size: members of a group run
  // in lockstep, so a small problem is not held to a much larger one's pace
  s += ":" + std::to_string((int)std::lround(std::log2((double)c->lev[c->L].N)));
  return s;
}

std::string plan_key(const std::vector<FuseJob*>& g) {
  std::string k;
  for (FuseJob* j : g) {
    char b[32];
    std::snprintf(b, sizeof b, "%p/", (void*)j->c);
    k += b + j->key + ";";
  }
  return k;
}

void drop_plan(hpcr_fuser::Plan& p) {
  if (p.exec) (void)hipGraphExecDestroy(p.exec);
  p.exec = nullptr;
}

// Issue the recorded sequences: position by position, one launch per kernel
// type present at that position.
void issue(hipStream_t st, std::vector<Recorder>& recs, hpcr_fuser::Plan& plan) {
  const size_t n = recs[0].recs.size();
  std::vector<const Rec*> ptr;
  std::vector<char> used(recs.size());
  for (size_t pos = 0; pos < n; pos++) {
    std::fill(used.begin(), used.end(), 0);
    for (size_t i = 0; i < recs.size(); i++) {
      if (used[i]) continue;
      const MultiFn fn = recs[i].recs[pos].fn;
      ptr.clear();
      for (size_t k = i; k < recs.size(); k++)
        if (!used[k] && recs[k].recs[pos].fn == fn) {
          used[k] = 1;
          ptr.push_back(&recs[k].recs[pos]);
        }
      fn(st, ptr.data(), (int)ptr.size());
      plan.launches += ((long long)ptr.size() + FUSE_GROUP_MAX - 1) / FUSE_GROUP_MAX;
      plan.problem_launches += (long long)ptr.size();
      plan.max_problems = std::max(plan.max_problems, std::min((int)ptr.size(), FUSE_GROUP_MAX));
    }
  }
}

// One batch of contexts with the same cycle signature (at most FUSE_GROUP_MAX).
// A job's key holds every value its recording passes to the kernels by value
// that can change between calls (the load, the first-step flag, the mixed
// form's buffer parity and coefficient epoch), so a cached graph of the same
// contexts and keys replays without recording again.
void run_group(hpcr_fuser* F, std::vector<FuseJob*>& g, hipStream_t st) {
  const std::string key = plan_key(g);
  F->tick++;
  if (F->graphs) {
    auto it = F->plans.find(key);
    if (it != F->plans.end()) {
      hpcr_fuser::Plan& p = it->second;
      hipError_t e = hipGraphLaunch(p.exec, st);
      for (size_t i = 0; i < g.size(); i++) {
        g[i]->iters = p.iters[i];
        if (e != hipSuccess) {
          g[i]->c->err = std::string("fused graph launch: ") + hipGetErrorString(e);
          g[i]->rc = HPCR_ERR_DEVICE;
        }
      }
      p.used = F->tick;
      F->st.graph_replays++;
      F->st.launches += p.launches;
      F->st.problem_launches += p.problem_launches;
      F->st.max_problems = std::max(F->st.max_problems, p.max_problems);
      return;
    }
  }
  std::vector<Recorder> recs(g.size());
  hpcr_fuser::Plan plan;
  for (size_t i = 0; i < g.size(); i++) {
    hpcr_ctx* c = g[i]->c;
    const long long it0 = c->smoother_iters;
    t_recorder = &recs[i];
    const int rc = g[i]->body();
    t_recorder = nullptr;
    g[i]->iters = c->smoother_iters - it0;
    c->smoother_iters = it0;
    g[i]->rc = rc;
    plan.members.push_back(c);
    plan.iters.push_back(g[i]->iters);
  }
  bool ok = true;
  for (size_t i = 0; i < g.size(); i++)
    if (g[i]->rc || recs[i].recs.size() != recs[0].recs.size() || recs[i].recs.empty()) ok = false;
  if (!ok) {
    // a failed recording, or sequences that do not line up: each member alone
    for (size_t i = 0; i < g.size(); i++) {
      if (g[i]->rc || recs[i].recs.empty()) continue;
      std::vector<Recorder> one(1);
      one[0].recs.swap(recs[i].recs);
      hpcr_fuser::Plan p1;
      issue(st, one, p1);
      F->st.launches += p1.launches;
      F->st.problem_launches += p1.problem_launches;
    }
    return;
  }
  hipGraph_t graph = nullptr;
  // capturing costs more than issuing eagerly: a batch is captured only when
  // its key comes back (the membership of batches changes often in serving)
  bool capture = false;
  if (F->graphs) {
    if (F->seen.size() > 4096) F->seen.clear();
    capture = F->seen[key]++ > 0;
  }
  if (capture && hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) capture = false;
  issue(st, recs, plan);
  if (capture) {
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (e == hipSuccess) e = hipGraphInstantiate(&plan.exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (e == hipSuccess) e = hipGraphLaunch(plan.exec, st);
    if (e != hipSuccess) {
      drop_plan(plan);
      for (FuseJob* j : g) {
        j->c->err = std::string("fused graph: ") + hipGetErrorString(e);
        j->rc = HPCR_ERR_DEVICE;
      }
      return;
    }
    F->st.graph_captures++;
    if (F->plans.size() >= FUSE_PLAN_MAX) {
      auto lru = F->plans.begin();
      for (auto it = F->plans.begin(); it != F->plans.end(); ++it)
        if (it->second.used < lru->second.used) lru = it;
      drop_plan(lru->second);
      F->plans.erase(lru);
    }
    plan.used = F->tick;
    F->plans[key] = plan;
  } else {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
      for (FuseJob* j : g) {
        j->c->err = std::string("fused launch: ") + hipGetErrorString(e);
        j->rc = HPCR_ERR_DEVICE;
      }
  }
  F->st.launches += plan.launches;
  F->st.problem_launches += plan.problem_launches;
  F->st.max_problems = std::max(F->st.max_problems, plan.max_problems);
}

// mark jobs as issued and wake their submitters
void release_jobs(hpcr_fuser* F, const std::vector<FuseJob*>& js) {
  std::lock_guard<std::mutex> lk(F->mu);
  for (FuseJob* j : js) j->done = true;
  F->cv_done.notify_all();
}

// A batch: jobs grouped by signature, each group on the next of the fuser's
// streams after its members' own work; a group's members are released as
// soon as its launches are issued.
void run_batch(hpcr_fuser* F, std::vector<FuseJob*>& jobs) {
  std::lock_guard<std::mutex> plk(F->plan_mu);
  std::vector<char> taken(jobs.size(), 0);
  for (size_t i = 0; i < jobs.size(); i++) {
    if (taken[i]) continue;
    std::vector<FuseJob*> g;
    for (size_t k = i; k < jobs.size() && (int)g.size() < FUSE_GROUP_MAX; k++)
      if (!taken[k] && jobs[k]->sig == jobs[i]->sig) {
        taken[k] = 1;
        g.push_back(jobs[k]);
      }
    // the group runs on its first member's own stream (a group of one runs
    // exactly where its context would run it)
    hipStream_t st = g[0]->c->stream;
    std::vector<FuseJob*> live;
    for (FuseJob* j : g) {
      if (j->c->stream != st && hipStreamWaitEvent(st, j->c->ev_in, 0) != hipSuccess) {
        j->c->err = "fuser: cannot order after the context's stream";
        j->rc = HPCR_ERR_DEVICE;
      } else {
        live.push_back(j);
      }
    }
    if (!live.empty()) run_group(F, live, st);
    for (FuseJob* j : g)
      if (hipEventRecord(j->c->ev_out, st) != hipSuccess && !j->rc) {
        j->c->err = "fuser: cannot record the group's end";
        j->rc = HPCR_ERR_DEVICE;
      }
    F->st.groups++;
    release_jobs(F, g);
  }
  F->st.batches++;
  F->st.cycles += (long long)jobs.size();
}

void fuser_loop(hpcr_fuser* F) {
  (void)hipSetDevice(F->device);
  std::unique_lock<std::mutex> lk(F->mu);
  for (;;) {
    F->cv_work.wait(lk, [&] { return F->stop || !F->pending.empty(); });
    if (F->pending.empty()) break;  // stop requested, nothing left
    const auto until = std::chrono::steady_clock::now() + std::chrono::microseconds(F->window_us);
    // wait (at most the window) only for in-loop contexts whose last cycle
    // could join a pending one's group
    auto partners_missing = [&] {
      for (const auto& kv : F->in_loop) {
        bool pending = false, partner = false;
        for (const FuseJob* j : F->pending) {
          if (j->c == kv.first) pending = true;
          if (j->sig == kv.second) partner = true;
        }
        if (partner && !pending) return true;
      }
      return false;
    };
    F->cv_work.wait_until(lk, until, [&] { return F->stop || !partners_missing(); });
    std::vector<FuseJob*> jobs;
    jobs.swap(F->pending);
    lk.unlock();
    run_batch(F, jobs);
    lk.lock();
  }
}

// whether another context inside a cycle loop of the same fuser has the
// signature sig (only then can a fused launch hold more than this context)
bool fuse_partner(hpcr_ctx* c, const std::string& sig) {
  hpcr_fuser* F = c->fuser;
  std::lock_guard<std::mutex> lk(F->mu);
  for (const auto& kv : F->in_loop)
    if (kv.first != c && kv.second == sig) return true;
  return false;
}

// hand a context's launch sequence (body) to its fuser; the caller waits
// until the batch is issued, and the context's stream then waits for the
// batch on the device
int fuse_submit(hpcr_ctx* c, std::string sig, std::string key, std::function<int()> body) {
  hpcr_fuser* F = c->fuser;
  HIPCHK(c, hipEventRecord(c->ev_in, c->stream));
  FuseJob j;
  j.c = c;
  j.sig = std::move(sig);
  j.key = std::move(key);
  j.body = std::move(body);
  {
    std::unique_lock<std::mutex> lk(F->mu);
    F->pending.push_back(&j);
    F->cv_work.notify_all();
    F->cv_done.wait(lk, [&] { return j.done; });
  }
  if (j.rc) return j.rc;
  HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_out, 0));
  c->smoother_iters += j.iters;
  return HPCR_OK;
}

int fused_cycle(hpcr_ctx* c, int kind, double Ut, int t) {
  char key[64];
  if (kind == 0)
    std::snprintf(key, sizeof key, "u/%a/%d", Ut, t == 0 ? 1 : 0);
  else
    std::snprintf(key, sizeof key, "d");
  int rc = fuse_submit(c, cycle_signature(c, kind), key,
                       [=] { return kind == 0 ? cycle_u(c, c->L, Ut, t) : cycle_d(c, c->L); });
  if (!rc) c->last_rr_level = c->L;
  return rc;
}

// marks an attached context as inside a solve's cycle loop
struct FuseLoop {
  hpcr_fuser* F;
  hpcr_ctx* c;
  FuseLoop(hpcr_ctx* c_, const std::string& sig) : F(c_->fuser), c(c_) {
    if (!F) return;
    std::lock_guard<std::mutex> lk(F->mu);
    F->in_loop[c] = sig;
  }
  ~FuseLoop() { release(); }
  void release() {
    if (!F) return;
    std::lock_guard<std::mutex> lk(F->mu);
    F->in_loop.erase(c);
    F->cv_work.notify_all();
    F = nullptr;
  }
};

// one mechanical cycle on the finest level, as a graph replay when enabled
// (kernel timing records events per launch, so it runs the cycle eagerly)
int run_cycle_u(hpcr_ctx* c, double Ut, int t) {
  if (c->fuser && fuse_partner(c, cycle_signature(c, 0))) return fused_cycle(c, 0, Ut, t);
  if (!c->graphs || c->timing) return cycle_u(c, c->L, Ut, t);
  const int t0 = (t == 0) ? 1 : 0;
  if (!c->g_u || c->g_u_Ut != Ut || c->g_u_t0 != t0) {
    const long long it0 = c->smoother_iters;
    int rc = capture_graph(c, &c->g_u, [&] { return cycle_u(c, c->L, Ut, t); });
    c->g_u_iters = c->smoother_iters - it0;
    c->smoother_iters = it0;
    if (rc) return rc;
    c->g_u_Ut = Ut;
    c->g_u_t0 = t0;
  }
  HIPCHK(c, hipGraphLaunch(c->g_u, c->stream));
  c->smoother_iters += c->g_u_iters;
  c->last_rr_level = c->L;
  return HPCR_OK;
}
int run_cycle_d(hpcr_ctx* c) {
  if (c->fuser && fuse_partner(c, cycle_signature(c, 1))) return fused_cycle(c, 1, 0.0, 0);
  if (!c->graphs || c->timing) return cycle_d(c, c->L);
  if (!c->g_d) {
    const long long it0 = c->smoother_iters;
    int rc = capture_graph(c, &c->g_d, [&] { return cycle_d(c, c->L); });
    c->g_d_iters = c->smoother_iters - it0;
    c->smoother_iters = it0;
    if (rc) return rc;
  }
  HIPCHK(c, hipGraphLaunch(c->g_d, c->stream));
  c->smoother_iters += c->g_d_iters;
  c->last_rr_level = c->L;
  return HPCR_OK;
}

// finest-level squared residual norm after a cycle
int finest_rr(hpcr_ctx* c, bool mech, double* out) {
  int rc;
  if (c->last_rr_level != c->L) {
    Lev& F = c->lev[c->L];
    launch_norm2(c->stream, (mech ? 3 : 1) * F.N, mech ? F.ru : F.rd, red(c));
  }
  if ((rc = read_scal(c))) return rc;
  *out = c->host_scal[S_RR];
  return HPCR_OK;
}

// u_entire (mechanical_displacement.cpp:25-115)
int u_entire(hpcr_ctx* c, double Ut, int t, double* f_norm, hpcr_stats* st) {
  int rc;
  const int L = c->L;
  Lev& F = c->lev[L];
  c->smoother_iters = 0;
  if ((rc = mech_setup(c))) return rc;
  if ((rc = mech_diag_all(c))) return rc;
  // Initial_residual_u (mechanical_pcg.cpp:193-220)
  launch_mech_resid(c->stream, F.g, mcoef(c, L), F.u, F.fu, F.ru, F.zu, F.invDu, 0, red(c));
  if ((rc = check_launch(c, "initial residual"))) return rc;
  if ((rc = read_scal(c))) return rc;
  double r_norm = c->host_scal[S_RR];
  if (st) {
    st->r_norm_initial = r_norm;
    st->cycles = 0;
  }
  double rel = 1.0;
  if (t >= 1 && *f_norm != 0) rel = r_norm / *f_norm;
  int cycles = 0;
  {
    FuseLoop in_loop(c, cycle_signature(c, 0));
    for (int j = 1; j <= c->par.mg_u.ncy; j++) {
      if (rel > c->par.tol_u) {
        if ((rc = run_cycle_u(c, Ut, t))) return rc;
        cycles++;
        if ((rc = finest_rr(c, true, &r_norm))) return rc;
        if (t == 0) *f_norm = c->host_scal[S_FF];
        rel = r_norm / *f_norm;
        if (st && cycles <= HPCR_HIST_MAX) st->history[cycles - 1] = r_norm;
      }
      if (rel <= c->par.tol_u) break;
    }
  }
  // externe_force_nodal (mechanical_right_hnd.cpp:18-38)
  launch_dirichlet_bd(c->stream, F.g, F.u, Ut, F.kplanes.data(), (int)F.kplanes.size());
  launch_mech_apply(c->stream, F.g, mcoef(c, L), F.u, F.w);
  launch_dirichlet_copy(c->stream, F.g, F.fu, F.w, F.kplanes.data(), (int)F.kplanes.size());
  if ((rc = check_launch(c, "externe_force_nodal"))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->diag_u_valid = false;
  if (st) {
    st->cycles = cycles;
    st->r_norm_final = r_norm;
    st->smoother_iters = c->smoother_iters;
  }
  return HPCR_OK;
}

// ---- mixed-precision form of u_entire ---------------------------------------
template <class P>
std::vector<MixLev<P>>& mixv(hpcr_ctx* c) {
  if constexpr (std::is_same<P, double>::value)
    return c->mix64;
  else
    return c->mix;
}

template <class P>
int mix_alloc(hpcr_ctx* c) {
  constexpr bool F64 = std::is_same<P, double>::value;
  if (!mixv<P>(c).empty()) return HPCR_OK;
  std::vector<MixLev<P>> mix(c->L + 1);
  int rc;
  for (int l = 0; l <= c->L; l++) {
    Lev& v = c->lev[l];
    const size_t N = v.N;
    MixLev<P>& m = mix[l];
    if ((rc = alloc_n(c, &m.x, 3 * N)) || (rc = alloc_n(c, &m.r, 3 * N))) return rc;
    if (l == 0) continue;
    if ((rc = alloc_n(c, &m.dA, 3 * N)) || (rc = alloc_n(c, &m.dB, 3 * N))) return rc;
    if constexpr (F64) {
      m.invD = v.invDu;
      if (l < c->L) {
        m.Ke = v.Ke;
        m.Ge = v.Ge;
      }
    } else {
      if ((rc = alloc_n(c, &m.invD, 3 * N))) return rc;
      if (l == c->L) {
        if ((rc = alloc_n(c, &m.Kn, N)) || (rc = alloc_n(c, &m.Gn, N)) || (rc = alloc_n(c, &m.dn, N))) return rc;
      } else {
        if ((rc = alloc_n(c, &m.Ke, N)) || (rc = alloc_n(c, &m.Ge, N))) return rc;
      }
    }
  }
  if (!c->q2 && (rc = alloc_n(c, &c->q2, 3 * c->lev[c->L].N))) return rc;
  mixv<P>(c) = std::move(mix);
  return HPCR_OK;
}

template <class P>
MixCoefT<P> mixcoef(hpcr_ctx* c, int l) {
  const MixLev<P>& m = mixv<P>(c)[l];
  MixCoefT<P> k{};
  k.fine = (l == c->L) ? 1 : 0;
  k.mfma = c->mix_mfma;
  k.lm = m.lm;
  if constexpr (std::is_same<P, double>::value) {
    k.Kn = c->lev[l].K;
    k.Gn = c->lev[l].G;
    k.dn = c->lev[l].d;
  } else {
    k.Kn = m.Kn;
    k.Gn = m.Gn;
    k.dn = m.dn;
  }
  k.Ke = m.Ke;
  k.Ge = m.Ge;
  k.m = mcoef(c, l);
  return k;
}

// the FP32 Gauss-point moduli of the matrix-core form, allocated on first use
// (64 B per finest element beyond the mixed solve's estimate)
int mix_lm(hpcr_ctx* c) {
  MixLev<float>& M = c->mix[c->L];
  Lev& F = c->lev[c->L];
  if (!M.lm) {
    int rc = alloc_n(c, &M.lm, 16 * F.N);
    if (rc) return rc;
  }
  launch_mix_lame(c->stream, F.g, F.K, F.G, F.d, F.flag, c->par.small_k, M.lm);
  return check_launch(c, "matrix-core moduli");
}

// preconditioner coefficients and diagonals (FP32 copies; the FP64 form
// reads the FAS path's own), power iteration and Chebyshev coefficients
template <class P>
int mix_setup(hpcr_ctx* c, const hpcr_mixed_params& mp) {
  constexpr bool F64 = std::is_same<P, double>::value;
  int rc;
  c->mix_epoch++;
  c->mix_mfma = mp.element_mfma;
  const int m = mp.cheb_degree;
  for (int l = 1; l <= c->L; l++) {
    Lev& v = c->lev[l];
    MixLev<P>& M = mixv<P>(c)[l];
    if constexpr (!F64) {
      launch_convert(c->stream, 3 * v.N, v.invDu, M.invD);
      if (l == c->L) {
        launch_convert(c->stream, v.N, v.K, M.Kn);
        launch_convert(c->stream, v.N, v.G, M.Gn);
        launch_convert(c->stream, v.N, v.d, M.dn);
        if (mp.element_mfma && (rc = mix_lm(c))) return rc;
      } else {
        launch_convert(c->stream, v.N, v.Ke, M.Ke);
        launch_convert(c->stream, v.N, v.Ge, M.Ge);
      }
    }
    // power iteration on D^-1 A: Rayleigh quotient of the last iterate
    P* a = M.dA;
    P* b = M.dB;
    launch_power_init(c->stream, v.g, a);
    for (int it = 0; it < mp.power_iters; it++) {
      launch_mix_power(c->stream, v.g, mixcoef<P>(c, l), (const P*)a, b, (const P*)M.invD, red(c));
      std::swap(a, b);
    }
    if ((rc = check_launch(c, "mixed set-up"))) return rc;
    if ((rc = read_scal(c))) return rc;
    M.lam = c->host_scal[S_LAM];
    if (!(M.lam > 0.0) || !std::isfinite(M.lam))
      return set_err(c, HPCR_ERR_STATE, "power iteration gave no positive eigenvalue estimate");
    const double hi = mp.cheb_upper * M.lam, lo = hi / mp.cheb_ratio;
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    M.cr0 = (P)(1.0 / theta);
    M.cd.assign(m, P(0));
    M.cr.assign(m, P(0));
    double rho_prev = 1.0 / sigma;
    for (int j = 1; j < m; j++) {
      const double rho = 1.0 / (2.0 * sigma - rho_prev);
      M.cd[j] = (P)(rho * rho_prev);
      M.cr[j] = (P)(2.0 * rho / delta);
      rho_prev = rho;
    }
  }
  return HPCR_OK;
}

// one preconditioner V-cycle on the residual held in mix[L].r (smoother
// started); the result is mix[L].x.  With qprev, x.qprev goes to scal[S_ZQ].
template <class P>
int mix_vcycle(hpcr_ctx* c, int m, int nu0, const double* qprev) {
  constexpr bool F64 = std::is_same<P, double>::value;
  const int L = c->L;
  std::vector<MixLev<P>>& mix = mixv<P>(c);
  for (int l = L; l >= 1; l--) {
    MixLev<P>& M = mix[l];
    const LevelGeom& g = c->lev[l].g;
    const MixCoefT<P> C = mixcoef<P>(c, l);
    P* sv = M.dA;
    P* ov = M.dB;
    for (int j = 1; j < m; j++) {
      launch_mix_cheb(c->stream, g, C, (const P*)sv, ov, M.r, M.x, (const P*)M.invD, M.cd[j], M.cr[j], nullptr,
                      red(c));
      std::swap(sv, ov);
    }
    launch_mix_res(c->stream, g, C, (const P*)sv, M.r);
    MixLev<P>& K = mix[l - 1];
    if (l - 1 >= 1)
      launch_mix_restrict(c->stream, c->lev[l - 1].g, g, (const P*)M.r, K.r, K.dA, K.x, (const P*)K.invD, K.cr0, 1);
    else
      launch_mix_restrict(c->stream, c->lev[0].g, g, (const P*)M.r, K.r, (P*)nullptr, K.x, (const P*)nullptr, P(0),
                          0);
  }
  {
    Lev& v0 = c->lev[0];
    if (!F64 && small_level(v0.g)) {
      if constexpr (!F64)
        launch_small_pcg_u_f32(c->stream, v0.g, v0.Ke, v0.Ge, mix[0].x, mix[0].r, v0.invDu, nu0,
                               c->scal + S_SCRATCH);
    } else {
      // the FP64 smoother of the FAS path on the coarsest level's own buffers
      // (zero start, homogeneous boundary values)
      launch_convert(c->stream, 3 * v0.N, (const P*)mix[0].r, v0.fu);
      launch_fill(c->stream, 3 * v0.N, v0.u, 0.0);
      int rc = pcg_u(c, 0, nu0, 0.0, 1);
      if (rc) return rc;
      launch_convert(c->stream, 3 * v0.N, (const double*)v0.u, mix[0].x);
    }
  }
  for (int l = 1; l <= L; l++) {
    MixLev<P>& M = mix[l];
    const LevelGeom& g = c->lev[l].g;
    const MixCoefT<P> C = mixcoef<P>(c, l);
    launch_mix_prolong(c->stream, g, c->lev[l - 1].g, M.x, M.dA, (const P*)mix[l - 1].x);
    P* sv = M.dA;
    P* ov = M.dB;
    launch_mix_cheb(c->stream, g, C, (const P*)sv, ov, M.r, M.x, (const P*)M.invD, P(0), M.cr0,
                    (l == L && m == 1) ? qprev : nullptr, red(c));
    std::swap(sv, ov);
    for (int j = 1; j < m; j++) {
      launch_mix_cheb(c->stream, g, C, (const P*)sv, ov, M.r, M.x, (const P*)M.invD, M.cd[j], M.cr[j],
                      (l == L && j == m - 1) ? qprev : nullptr, red(c));
      std::swap(sv, ov);
    }
  }
  return check_launch(c, "mixed V-cycle");
}

template <class P>
int u_entire_mixed_p(hpcr_ctx* c, double Ut, int t, double* f_norm, const hpcr_mixed_params& mp, hpcr_stats* st) {
  int rc;
  const int L = c->L;
  Lev& F = c->lev[L];
  if (L < 1) return set_err(c, HPCR_ERR_ARG, "mixed precision needs at least two levels");
  if (mp.coarse_iters < 0) return set_err(c, HPCR_ERR_ARG, "coarse_iters must be >= 0");
  if (mp.cheb_degree < 1 || mp.cheb_degree > 8 || mp.power_iters < 1 || !(mp.cheb_ratio > 1.0) ||
      !(mp.cheb_upper > 0.0) || mp.max_iter < 0)
    return set_err(c, HPCR_ERR_ARG, "bad mixed-precision parameters");
  if ((rc = mix_alloc<P>(c))) return rc;
  c->smoother_iters = 0;
  if ((rc = mech_setup(c))) return rc;
  if ((rc = mech_diag_all(c))) return rc;
  if ((rc = mix_setup<P>(c, mp))) return rc;
  MixLev<P>& MF = mixv<P>(c)[L];
  const int t0 = (t == 0) ? 1 : 0;
  // residual from u with its boundary values; at t == 0 also f := A u on
  // Dirichlet DOFs and f_norm = |f|^2, as the finest smoother does
  auto true_residual = [&](double* rr) -> int {
    launch_dirichlet_bd(c->stream, F.g, F.u, Ut, F.kplanes.data(), (int)F.kplanes.size());
    launch_mech_resid(c->stream, F.g, mcoef(c, L), F.u, F.fu, F.ru, F.zu, F.invDu, t0, red(c));
    int e = check_launch(c, "mixed residual");
    if (e) return e;
    if ((e = read_scal(c))) return e;
    *rr = c->host_scal[S_RR];
    if (t0) *f_norm = c->host_scal[S_FF];
    launch_mix_start(c->stream, F.g, F.ru, MF.r, MF.dA, MF.x, (const P*)MF.invD, MF.cr0);
    return check_launch(c, "mixed start");
  };
  double rr = 0.0;
  if ((rc = true_residual(&rr))) return rc;
  if (st) st->r_norm_initial = rr;
  // relative residual as u_entire tests it (mechanical_displacement.cpp:84-102):
  // with f_norm == 0 (e.g. a zero load) a zero residual is converged and a
  // nonzero one never is, so the solve runs to its iteration cap like the
  // reference's cycles do
  auto relres = [&](double r) {
    return (*f_norm != 0.0) ? r / *f_norm : (r == 0.0 ? 0.0 : HUGE_VAL);
  };
  double rel = (rr == 0.0) ? 0.0 : ((*f_norm != 0.0) ? rr / *f_norm : 1.0);
  double* pb[2] = {F.p, F.p2};
  double* qb[2] = {F.w, c->q2};
  int it = 0;
  const int m = mp.cheb_degree;
  const int nu0 = mp.coarse_iters > 0 ? mp.coarse_iters : c->par.mg_u.nu0;
  // Pass B leaves u += alpha p to the next pass A; `lag` says that the update
  // of the last iteration (direction pb[it & 1]) is still to be applied,
  // which `flush` does where u is read before another pass A runs.
  bool lag = false;
  auto flush = [&]() {
    if (lag) launch_fcg_flush(c->stream, F.g, F.u, pb[it & 1], red(c));
    lag = false;
  };
  // one outer iteration's launches: preconditioner V-cycle, FCG passes A and B
  auto iteration = [&]() -> int {
    int e = mix_vcycle<P>(c, m, nu0, it > 0 ? qb[it & 1] : nullptr);
    if (e) return e;
    launch_fcg_a(c->stream, F.g, mcoef(c, L), (const P*)MF.x, pb[it & 1], pb[(it + 1) & 1], qb[(it + 1) & 1], F.ru,
                 it == 0, lag ? F.u : nullptr, red(c));
    launch_fcg_b(c->stream, F.g, F.u, F.ru, pb[(it + 1) & 1], qb[(it + 1) & 1], MF.r, MF.dA, MF.x, (const P*)MF.invD,
                 MF.cr0, true, red(c));
    return check_launch(c, "flexible CG");
  };
  const std::string msig =
      c->fuser ? cycle_signature(c, 2) + "/m" + std::to_string(m) + "/c" + std::to_string(nu0) +
                     (std::is_same<P, double>::value ? "/p64" : "")
               : std::string();
  FuseLoop in_loop(c, msig);
  // at t == 0 the reference always runs a first cycle; with an exactly zero
  // residual there is no direction to search (and p.q = 0)
  while (it < mp.max_iter && (rel > mp.tol || (t0 && it == 0 && rr != 0.0))) {
    if (c->fuser && fuse_partner(c, msig)) {
      char key[64];
      std::snprintf(key, sizeof key, "x/%d/%d/%d/%llu", it == 0 ? 1 : 0, it & 1, lag ? 1 : 0, c->mix_epoch);
      rc = fuse_submit(c, msig, key, iteration);
    } else {
      rc = iteration();
    }
    if (rc) return rc;
    lag = true;
    if ((rc = read_scal(c))) return rc;
    rr = c->host_scal[S_RR];
    it++;
    if (!std::isfinite(rr)) return set_err(c, HPCR_ERR_STATE, "mixed-precision iteration diverged");
    rel = relres(rr);
    const bool replace = (mp.replace_every > 0 && it % mp.replace_every == 0) || rel <= mp.tol;
    if (replace) {
      flush();
      if ((rc = true_residual(&rr))) return rc;
      rel = relres(rr);
    }
    if (st && it <= HPCR_HIST_MAX) st->history[it - 1] = rr;
  }
  in_loop.release();
  flush();
  if ((rc = check_launch(c, "flexible CG update"))) return rc;
  // externe_force_nodal (mechanical_right_hnd.cpp:18-38)
  launch_dirichlet_bd(c->stream, F.g, F.u, Ut, F.kplanes.data(), (int)F.kplanes.size());
  launch_mech_apply(c->stream, F.g, mcoef(c, L), F.u, F.w);
  launch_dirichlet_copy(c->stream, F.g, F.fu, F.w, F.kplanes.data(), (int)F.kplanes.size());
  if ((rc = check_launch(c, "externe_force_nodal"))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->diag_u_valid = false;
  c->last_rr_level = -1;
  if (st) {
    st->cycles = it;
    st->r_norm_final = rr;
    st->smoother_iters = 0;
  }
  return HPCR_OK;
}

// cheb_degree 0: the smoother setting from the finest level's bulk-modulus
// contrast max K / min K (K > 0).  Near-void phases (contrast above
// MIX_AUTO_CONTRAST, e.g. pores at 1e-6 E) make the largest eigenvalues of
// D^-1 A spread, and a wider, higher-degree polynomial pays for its extra
// passes; moderate contrast keeps the cheap degree-2 setting.
constexpr double MIX_AUTO_CONTRAST = 1e3;
int mixed_resolve(hpcr_ctx* c, hpcr_mixed_params& mp, double* contrast) {
  if (mp.cheb_degree != 0) return HPCR_OK;
  Lev& F = c->lev[c->L];
  launch_contrast(c->stream, (size_t)F.g.nx * F.g.ny * F.g.nz, F.K, red(c));
  int rc = check_launch(c, "modulus contrast");
  if (rc || (rc = read_scal(c))) return rc;
  const double k = c->host_scal[S_RED] * c->host_scal[S_SCRATCH];
  if (contrast) *contrast = k;
  if (k > MIX_AUTO_CONTRAST) {
    mp.cheb_degree = 4;
    mp.cheb_ratio = 12.0;
    mp.cheb_upper = 1.3;
  } else {
    mp.cheb_degree = 2;
    mp.cheb_ratio = 6.0;
    mp.cheb_upper = 1.1;
  }
  return HPCR_OK;
}

int u_entire_mixed(hpcr_ctx* c, double Ut, int t, double* f_norm, const hpcr_mixed_params& mp_in, hpcr_stats* st) {
  hpcr_mixed_params mp = mp_in;
  if (mp.precond_fp64 != 0 && mp.precond_fp64 != 1) return set_err(c, HPCR_ERR_ARG, "precond_fp64 must be 0 or 1");
  if (mp.element_mfma != 0 && mp.element_mfma != 1) return set_err(c, HPCR_ERR_ARG, "element_mfma must be 0 or 1");
  if (mp.element_mfma && mp.precond_fp64)
    return set_err(c, HPCR_ERR_ARG, "the matrix-core element contraction is an FP32 V-cycle form");
  int rc = mixed_resolve(c, mp, nullptr);
  if (rc) return rc;
  return mp.precond_fp64 ? u_entire_mixed_p<double>(c, Ut, t, f_norm, mp, st)
                         : u_entire_mixed_p<float>(c, Ut, t, f_norm, mp, st);
}

// ---- phase field ------------------------------------------------------------
int pf_diag(hpcr_ctx* c, int l, double* diag_out) {
  launch_pf_diag(c->stream, c->lev[l].g, pcoef(c, l), diag_out, c->lev[l].invDd, red(c));
  return check_launch(c, "pf_diag");
}

// pcg_d (phase_field_pcg.cpp:133-214)
int pcg_d(hpcr_ctx* c, int l, int NB) {
  Lev& v = c->lev[l];
  int rc;
  if (!c->diag_d_valid) {
    if ((rc = pf_diag(c, l, nullptr))) return rc;
    if ((rc = read_scal(c))) return rc;
    if (c->host_scal[S_ZERO] != 0.0)
      return set_err(c, HPCR_ERR_ZERO_DIAGONAL, "Division by zero in grid /= operation (phase-field diagonal)");
  }
  if (l < c->L && small_level(v.g)) {
    launch_small_pcg_d(c->stream, v.g, v.H, v.gc, c->par.lc, v.d, v.fd, v.invDd, v.rd, v.zd, (NB & 1) ? v.pd2 : v.pd,
                       v.wd, NB, c->scal);
    c->smoother_iters += NB;
    c->last_rr_level = l;
    return check_launch(c, "pcg_d (small level)");
  }
  launch_pf_resid(c->stream, v.g, pcoef(c, l), v.d, v.fd, v.rd, v.zd, v.invDd, red(c));
  double* pb[CG_GROUP] = {v.pd, v.pd2, v.pd3, v.pd4};
  const PGroup P{{v.pd, v.pd2, v.pd3, v.pd4}};
  for (int i = 0; i < NB; i++) {
    const int m = i % CG_GROUP;
    {
      TBracket tb(c, HPCR_K_PF_PCG_A, l);
      launch_pf_pcg_a(c->stream, v.g, pcoef(c, l), v.zd, pb[(m + CG_GROUP - 1) % CG_GROUP], pb[m], v.wd, i == 0,
                      red(c));
    }
    {
      // d += alpha_i p_i waits for the pass B that closes the group
      const int np = (m == CG_GROUP - 1 || i == NB - 1) ? m + 1 : 0;
      TBracket tb(c, HPCR_K_PF_PCG_B, l);
      launch_pf_pcg_b(c->stream, v.g, v.d, v.rd, v.zd, P, v.wd, v.invDd, np, m, red(c));
    }
    c->smoother_iters++;
  }
  c->last_rr_level = l;
  return check_launch(c, "pcg_d");
}

// cycle_d (phase_field_multigrid.cpp:18-44)
int cycle_d(hpcr_ctx* c, int l) {
  int rc;
  if (l == 0) return pcg_d(c, 0, c->par.mg_d.nu0);
  if ((rc = pcg_d(c, l, c->par.mg_d.nu1))) return rc;
  Lev& f = c->lev[l];
  Lev& k = c->lev[l - 1];
  // coarsen_d + coarsen_f_d (multigrid_intergrid.cpp:92-103, 345-371)
  launch_inject(c->stream, k.g, f.g, 1, f.d, k.d, k.dold, 0, 0.0);
  launch_pf_apply(c->stream, k.g, pcoef(c, l - 1), k.d, k.wd);
  launch_restrict(c->stream, k.g, f.g, 1, f.rd, k.wd, k.fd);
  for (int j = 1; j <= c->par.mg_d.gamma; j++)
    if ((rc = cycle_d(c, l - 1))) return rc;
  // refine_d (multigrid_intergrid.cpp:523-533)
  launch_prolong(c->stream, f.g, k.g, 1, f.d, k.d, k.dold, 0, 0.0, k.wd);
  for (int i = 1; i <= c->par.mg_d.nu2; i++)
    if ((rc = pcg_d(c, l, c->par.mg_d.nu2))) return rc;
  return check_launch(c, "cycle_d");
}

int pf_diag_all(hpcr_ctx* c) {
  for (int l = 0; l <= c->L; l++) {
    int rc = pf_diag(c, l, nullptr);
    if (rc) return rc;
    if ((rc = read_scal(c))) return rc;
    if (c->host_scal[S_ZERO] != 0.0)
      return set_err(c, HPCR_ERR_ZERO_DIAGONAL, "Division by zero in grid /= operation (phase-field diagonal)");
  }
  c->diag_d_valid = true;
  return HPCR_OK;
}

// phase_field (phase_field.cpp:15-57)
int phase_field(hpcr_ctx* c, int t, double* fd_norm, hpcr_stats* st) {
  (void)t;
  int rc;
  const int L = c->L;
  Lev& F = c->lev[L];
  c->smoother_iters = 0;
  if ((rc = pf_diag_all(c))) return rc;
  launch_pf_resid(c->stream, F.g, pcoef(c, L), F.d, F.fd, F.rd, F.zd, F.invDd, red(c));
  if ((rc = check_launch(c, "initial residual"))) return rc;
  if ((rc = read_scal(c))) return rc;
  double r_norm = c->host_scal[S_RR];
  if (st) st->r_norm_initial = r_norm;
  double sres = 1.0;
  int cycles = 0;
  {
    FuseLoop in_loop(c, cycle_signature(c, 1));
    for (int j = 1; j <= c->par.mg_d.ncy; j++) {
      if (r_norm > c->par.tol_d_abs) {
        if (sres > c->par.tol_d_rel) {
          if ((rc = run_cycle_d(c))) return rc;
          cycles++;
        }
        if ((rc = finest_rr(c, false, &r_norm))) return rc;
        *fd_norm = c->host_scal[S_FF];
        sres = r_norm / *fd_norm;
        if (st && cycles >= 1 && cycles <= HPCR_HIST_MAX) st->history[cycles - 1] = r_norm;
        if (sres <= c->par.tol_d_rel) break;
      } else {
        break;
      }
    }
  }
  c->diag_d_valid = false;
  if (st) {
    st->cycles = cycles;
    st->r_norm_final = r_norm;
    st->smoother_iters = c->smoother_iters;
  }
  return HPCR_OK;
}

// ---- field access ----------------------------------------------------------
bool field_vec(int f) { return f == HPCR_U || f == HPCR_FU || f == HPCR_RU || f == HPCR_UOLD || f == HPCR_P || f == HPCR_W; }

double* field_ptr(hpcr_ctx* c, int f, int l) {
  Lev& v = c->lev[l];
  switch (f) {
    case HPCR_U: return v.u;
    case HPCR_FU: return v.fu;
    case HPCR_RU: return v.ru;
    case HPCR_UOLD: return v.uold;
    case HPCR_P: return v.p;
    case HPCR_W: return v.w;
    case HPCR_D: return v.d;
    case HPCR_FD: return v.fd;
    case HPCR_RD: return v.rd;
    case HPCR_DOLD: return v.dold;
    case HPCR_H: return v.H;
    case HPCR_K: return v.K;
    case HPCR_G: return v.G;
    case HPCR_GC: return v.gc;
    case HPCR_KE: return v.Ke;
    case HPCR_GE: return v.Ge;
    case HPCR_DPREV: return v.dprev;
    case HPCR_PD: return v.pd;
    case HPCR_WD: return v.wd;
    default: return nullptr;
  }
}

int check_fl(hpcr_ctx* c, int f, int l, int comp) {
  if (!c) return HPCR_ERR_ARG;
  if (l < 0 || l > c->L) return set_err(c, HPCR_ERR_ARG, "level out of range");
  if (f < 0 || f >= HPCR_FIELD_COUNT) return set_err(c, HPCR_ERR_ARG, "unknown field");
  if (comp < 0 || comp >= (field_vec(f) ? 3 : 1)) return set_err(c, HPCR_ERR_ARG, "component out of range");
  if (f != HPCR_FLAG && !field_ptr(c, f, l)) return set_err(c, HPCR_ERR_ARG, "field not held on this level");
  if (f == HPCR_FLAG && !c->lev[l].flag) return set_err(c, HPCR_ERR_ARG, "field not held on this level");
  return HPCR_OK;
}

int build_levels(hpcr_ctx* c) {
  const hpcr_geometry& G = c->geo;
  const int L = G.maxlevel;
  c->L = L;
  c->lev.resize(L + 1);
  // initialize (initialization.cpp:19-134), one rank per axis
  double hx = (G.xb - G.xa) / G.element_nb[0];
  double hy = (G.yb - G.ya) / G.element_nb[1];
  double hz = (G.zb - G.za) / G.element_nb[2];
  long ii = G.element_nb[0], jj = G.element_nb[1], kk = G.element_nb[2];
  size_t max_tile_blocks = 1;
  for (int l = 0; l <= L; l++) {
    Lev& v = c->lev[l];
    v.g.nx = (int)ii + 1;
    v.g.ny = (int)jj + 1;
    v.g.nz = (int)kk + 1;
    v.g.hx = hx;
    v.g.hy = hy;
    v.g.hz = hz;
    v.N = (size_t)v.g.nx * v.g.ny * v.g.nz;
    v.g.zc = choose_zchunk(v.g.nx, v.g.ny, v.g.nz);
    hx *= 0.5; hy *= 0.5; hz *= 0.5;
    ii *= 2; jj *= 2; kk *= 2;
    size_t tb = (size_t)tile_blocks(v.g.nx, v.g.ny, v.g.nz, v.g.zc);
    if (tb > max_tile_blocks) max_tile_blocks = tb;
    if (l == L && split_pass_a(v.g.nx, v.g.ny)) {
      v.sw = split_dims(v.g.nx, v.g.ny, v.g.nz);
      tb = (size_t)split_blocks(v.g, v.sw);
      if (tb > max_tile_blocks) max_tile_blocks = tb;
    }
  }
  int rc;
  for (int l = 0; l <= L; l++) {
    Lev& v = c->lev[l];
    const size_t N = v.N;
    double** vec3[] = {&v.u, &v.fu, &v.ru, &v.p, &v.p2, &v.p3, &v.p4, &v.w, &v.invDu, &v.zu};
    for (double** pp : vec3)
      if ((rc = alloc_n(c, pp, 3 * N))) return rc;
    if (l < L && (rc = alloc_n(c, &v.uold, 3 * N))) return rc;
    double** sc[] = {&v.d,  &v.fd, &v.rd, &v.pd, &v.pd2, &v.pd3, &v.pd4, &v.wd,
                     &v.invDd, &v.zd, &v.H, &v.K, &v.G, &v.gc, &v.Ke, &v.Ge};
    for (double** pp : sc)
      if ((rc = alloc_n(c, pp, N))) return rc;
    if (l < L && (rc = alloc_n(c, &v.dold, N))) return rc;
    if (l == L) {
      if ((rc = alloc_n(c, &v.flag, N))) return rc;
      if ((rc = alloc_n(c, &v.dprev, N))) return rc;
      if (v.sw.ntx > 0 && (rc = alloc_n(c, &v.sw.ws, v.sw.count()))) return rc;
    }
    // Dirichlet axis flags by coordinate equality (boundary_conditions.cpp:28-66)
    std::vector<unsigned char> fx(v.g.nx), fy(v.g.ny), fz(v.g.nz);
    const double xmid = 0.5 * (G.xa + G.xb), ymid = 0.5 * (G.ya + G.yb);
    for (int i = 0; i < v.g.nx; i++) {
      const double x = coord(G.xa, v.g.hx, i);
      fx[i] = (x == xmid ? 1 : 0) | (x == G.xb ? 2 : 0);
    }
    for (int j = 0; j < v.g.ny; j++) {
      const double y = coord(G.ya, v.g.hy, j);
      fy[j] = (y == G.ya ? 1 : 0) | (y == ymid ? 2 : 0);
    }
    for (int k = 0; k < v.g.nz; k++) {
      const double z = coord(G.za, v.g.hz, k);
      fz[k] = (z == G.za ? 1 : 0) | (z == G.zb ? 2 : 0);
      if (fz[k]) v.kplanes.push_back(k);
    }
    if (v.kplanes.size() > 4) return set_err(c, HPCR_ERR_ARG, "more than 4 Dirichlet z-planes on one level");
    if ((rc = alloc_n(c, &v.fx, v.g.nx)) || (rc = alloc_n(c, &v.fy, v.g.ny)) || (rc = alloc_n(c, &v.fz, v.g.nz)))
      return rc;
    HIPCHK(c, hipMemcpyAsync(v.fx, fx.data(), fx.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(v.fy, fy.data(), fy.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(v.fz, fz.data(), fz.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));  // the host vectors end with this iteration
    v.g.fx = v.fx;
    v.g.fy = v.fy;
    v.g.fz = v.fz;
  }
  {
    long pb = 4096;
    if (max_tile_blocks > pb) pb = max_tile_blocks;
    for (const Lev& v : c->lev)
      if (passb_blocks(v.N) > pb) pb = passb_blocks(v.N);
    c->partial_cap = (size_t)pb * 4;
  }
  if ((rc = alloc_n(c, &c->partials, c->partial_cap))) return rc;
  if ((rc = alloc_n(c, &c->counter, 1))) return rc;
  if ((rc = alloc_n(c, &c->scal, S_COUNT))) return rc;
  HIPCHK(c, hipHostMalloc((void**)&c->host_scal, S_COUNT * sizeof(double), hipHostMallocDefault));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HPCR_OK;
}

// device bytes that build_levels (and, with mixed, mix_alloc) allocate
// mixed: 0 none, 1 with the FP32 preconditioner's buffers, 2 with the FP64 one's
long long estimate_bytes(const hpcr_geometry& G, int mixed) {
  const int L = G.maxlevel;
  long long total = 0, pb = 4096;
  long ii = G.element_nb[0], jj = G.element_nb[1], kk = G.element_nb[2];
  for (int l = 0; l <= L; l++) {
    const long long nx = ii + 1, ny = jj + 1, nz = kk + 1, N = nx * ny * nz;
    total += 10 * 3 * N * 8 + (l < L ? 3 * N * 8 : 0) + 16 * N * 8 + (l < L ? N * 8 : 0);
    if (l == L) total += N + N * 8;
    total += nx + ny + nz;
    const int zc = choose_zchunk((int)nx, (int)ny, (int)nz);
    const long tb = tile_blocks((int)nx, (int)ny, (int)nz, zc);
    if (tb > pb) pb = tb;
    if (l == L && split_pass_a((int)nx, (int)ny)) {
      const SplitW S = split_dims((int)nx, (int)ny, (int)nz);
      total += (long long)(S.count() * 8);
      const long sb = (long)S.ntx * S.nty * ((nz + S.zc - 1) / S.zc);
      if (sb > pb) pb = sb;
    }
    if (passb_blocks((size_t)N) > pb) pb = passb_blocks((size_t)N);
    if (mixed == 1) {
      total += 2 * 3 * N * 4;
      if (l > 0) total += 3 * 3 * N * 4 + (l == L ? 3 * N * 4 : 2 * N * 4);
      if (l == L) total += 3 * N * 8;
    } else if (mixed == 2) {
      total += 2 * 3 * N * 8;
      if (l > 0) total += 2 * 3 * N * 8;
      if (l == L) total += 3 * N * 8;
    }
    ii *= 2;
    jj *= 2;
    kk *= 2;
  }
  total += pb * 4 * 8 + 4 + S_COUNT * 8;
  return total;
}

// detach a context from its fuser: cached batch graphs that hold it go
void fuser_detach(hpcr_ctx* c) {
  hpcr_fuser* F = c->fuser;
  if (!F) return;
  {
    std::lock_guard<std::mutex> plk(F->plan_mu);
    for (auto it = F->plans.begin(); it != F->plans.end();) {
      if (std::find(it->second.members.begin(), it->second.members.end(), c) != it->second.members.end()) {
        // (an executable graph destroyed while a replay runs is freed when
        // the replay completes; the context's own stream waits for every
        // fused batch it was in, and destroy() synchronises it before freeing)
        drop_plan(it->second);
        it = F->plans.erase(it);
      } else {
        ++it;
      }
    }
  }
  {
    std::lock_guard<std::mutex> lk(F->mu);
    F->attached--;
  }
  c->fuser = nullptr;
}

void destroy(hpcr_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  fuser_detach(c);
  if (c->ev_in) (void)hipEventDestroy(c->ev_in);
  if (c->ev_out) (void)hipEventDestroy(c->ev_out);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& r : c->trec) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  if (c->g_u) (void)hipGraphExecDestroy(c->g_u);
  if (c->g_d) (void)hipGraphExecDestroy(c->g_d);
  if (c->stream) {
    for (void* p : c->allocs) (void)hipFreeAsync(p, c->stream);
    (void)hipStreamSynchronize(c->stream);
  }
  if (c->host_scal) (void)hipHostFree(c->host_scal);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

}  // namespace

// -----------------------------------------------------------------------------
// Device scratch of hpcr_set_image_noalloc: the material table, the
// out-of-table flag and a staging area through which the image streams in
// chunks of NOALLOC_STAGE bytes.  One copy per device (each at its own
// address), held until the process ends.  Every call on a device shares that
// device's copy, so calls that overlap on one device race.
// -----------------------------------------------------------------------------
namespace {
constexpr size_t NOALLOC_STAGE = (size_t)4 << 20;
constexpr size_t NOALLOC_TAB_OFF = 0;
constexpr size_t NOALLOC_BAD_OFF = (sizeof(MatTable) + 255) / 256 * 256;
constexpr size_t NOALLOC_IMG_OFF = NOALLOC_BAD_OFF + 256;
constexpr size_t NOALLOC_BYTES = NOALLOC_IMG_OFF + NOALLOC_STAGE;
}  // namespace
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// The device address of this device's scratch, looked up once per device
// with that device current.  The host-side name `scratch` is a stand-in, not
// a device address: host code uses only the pointer returned here.
int scratch_of(hpcr_ctx* c, unsigned char** out) {
  if (c->device < 0 || c->device >= NOALLOC_MAX_DEVICES) return set_err(c, HPCR_ERR_DEVICE, "device index too large");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[c->device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_err(c, HPCR_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[c->device] = (unsigned char*)p;
  }
  *out = g_scratch_ptr[c->device];
  return HPCR_OK;
}

__global__ void zero_flag_kernel(int* f) {
  if (threadIdx.x == 0) f[0] = 0;
}
}  // namespace

// =============================================================================
// C-ABI
// =============================================================================
extern "C" {

int hpcr_device_count(int* n) {
  if (!n) return HPCR_ERR_ARG;
  *n = 0;
  hipError_t e = hipGetDeviceCount(n);
  if (e != hipSuccess) {
    *n = 0;
    return HPCR_ERR_DEVICE;
  }
  return HPCR_OK;
}

const char* hpcr_last_error(const hpcr_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int hpcr_ctx_create(const hpcr_geometry* geo, const hpcr_params* par, int device, hpcr_ctx** out) {
  g_create_error.clear();
  if (!geo || !par || !out) {
    g_create_error = "null argument";
    return HPCR_ERR_ARG;
  }
  *out = nullptr;
  if (geo->maxlevel < 0 || geo->maxlevel > 12 || geo->element_nb[0] < 1 || geo->element_nb[1] < 1 ||
      geo->element_nb[2] < 1) {
    g_create_error = "bad geometry";
    return HPCR_ERR_ARG;
  }
  if (par->mg_u.ncy < 0 || par->mg_d.ncy < 0 || par->mg_u.gamma < 1 || par->mg_d.gamma < 1) {
    g_create_error = "bad multigrid parameters";
    return HPCR_ERR_ARG;
  }
  {
    // the kernels address a node of one field component by a 32-bit byte offset
    long long n = 1;
    for (int d = 0; d < 3; d++) n *= (long long)geo->element_nb[d] * (1LL << geo->maxlevel) + 1;
    if (n > MAX_LEVEL_NODES) {
      g_create_error = "finest level above 2^29 nodes";
      return HPCR_ERR_ARG;
    }
  }
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    g_create_error = "no HIP device";
    return HPCR_ERR_DEVICE;
  }
  if (device < 0 || device >= n) {
    g_create_error = "device index out of range";
    return HPCR_ERR_ARG;
  }
  hpcr_ctx* c = new hpcr_ctx();
  c->device = device;
  c->geo = *geo;
  c->par = *par;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    g_create_error = "cannot create stream";
    destroy(c);
    return HPCR_ERR_DEVICE;
  }
  int rc = build_levels(c);
  if (rc) {
    g_create_error = c->err;
    destroy(c);
    return rc;
  }
  *out = c;
  return HPCR_OK;
}

int hpcr_ctx_destroy(hpcr_ctx* ctx) {
  destroy(ctx);
  return HPCR_OK;
}

int hpcr_level_dims(const hpcr_ctx* c, int l, int dims[3]) {
  if (!c || !dims || l < 0 || l > c->L) return HPCR_ERR_ARG;
  dims[0] = c->lev[l].g.nx;
  dims[1] = c->lev[l].g.ny;
  dims[2] = c->lev[l].g.nz;
  return HPCR_OK;
}
int hpcr_level_h(const hpcr_ctx* c, int l, double h[3]) {
  if (!c || !h || l < 0 || l > c->L) return HPCR_ERR_ARG;
  h[0] = c->lev[l].g.hx;
  h[1] = c->lev[l].g.hy;
  h[2] = c->lev[l].g.hz;
  return HPCR_OK;
}
long long hpcr_device_bytes(const hpcr_ctx* c) { return c ? c->bytes : 0; }

int hpcr_synchronize(hpcr_ctx* c) {
  if (!c) return HPCR_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HPCR_OK;
}

int hpcr_set_field(hpcr_ctx* c, int f, int l, int comp, const double* host) {
  int rc = check_fl(c, f, l, comp);
  if (rc) return rc;
  if (!host) return set_err(c, HPCR_ERR_ARG, "null host pointer");
  (void)hipSetDevice(c->device);
  Lev& v = c->lev[l];
  if (f == HPCR_FLAG) {
    HIPCHK(c, hipMemcpyAsync(v.wd, host, v.N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    launch_double_to_flag(c->stream, v.N, v.wd, v.flag);
  } else {
    double* p = field_ptr(c, f, l) + (size_t)comp * v.N;
    HIPCHK(c, hipMemcpyAsync(p, host, v.N * sizeof(double), hipMemcpyHostToDevice, c->stream));
  }
  if (f == HPCR_K || f == HPCR_G || f == HPCR_D || f == HPCR_FLAG || f == HPCR_KE || f == HPCR_GE)
    c->diag_u_valid = false;
  if (f == HPCR_H || f == HPCR_GC) c->diag_d_valid = false;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HPCR_OK;
}

int hpcr_get_field(hpcr_ctx* c, int f, int l, int comp, double* host) {
  int rc = check_fl(c, f, l, comp);
  if (rc) return rc;
  if (!host) return set_err(c, HPCR_ERR_ARG, "null host pointer");
  (void)hipSetDevice(c->device);
  Lev& v = c->lev[l];
  if (f == HPCR_FLAG) {
    // borrow the scalar work field of the phase field as staging
    launch_flag_to_double(c->stream, v.N, v.flag, v.wd);
    HIPCHK(c, hipMemcpyAsync(host, v.wd, v.N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  } else {
    const double* p = field_ptr(c, f, l) + (size_t)comp * v.N;
    HIPCHK(c, hipMemcpyAsync(host, p, v.N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HPCR_OK;
}

int hpcr_fill_field(hpcr_ctx* c, int f, int l, int comp, double value) {
  int rc = check_fl(c, f, l, comp);
  if (rc) return rc;
  if (f == HPCR_FLAG) return set_err(c, HPCR_ERR_ARG, "fill of the flag field");
  (void)hipSetDevice(c->device);
  Lev& v = c->lev[l];
  launch_fill(c->stream, v.N, field_ptr(c, f, l) + (size_t)comp * v.N, value);
  c->diag_u_valid = c->diag_d_valid = false;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_launch(c, "fill");
}

int hpcr_coarse_vr(hpcr_ctx* c, int f) {
  if (!c) return HPCR_ERR_ARG;
  if (f != HPCR_K && f != HPCR_G && f != HPCR_GC) return set_err(c, HPCR_ERR_ARG, "coarse_vr takes K, G or GC");
  (void)hipSetDevice(c->device);
  for (int l = c->L; l > 0; l--)
    launch_nodal_vr(c->stream, c->lev[l - 1].g, c->lev[l].g, field_ptr(c, f, l), field_ptr(c, f, l - 1));
  c->diag_u_valid = c->diag_d_valid = false;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_launch(c, "coarse_vr");
}

int hpcr_set_reduction_fence(hpcr_ctx* c, int on) {
  if (!c) return HPCR_ERR_ARG;
  c->ticket_fence = on ? 1 : 0;
  return HPCR_OK;
}

int hpcr_set_graphs(hpcr_ctx* c, int on) {
  if (!c) return HPCR_ERR_ARG;
  c->graphs = on != 0;
  return HPCR_OK;
}

int hpcr_fuser_create(int device, int window_us, hpcr_fuser** out) {
  if (!out || window_us < 0) return HPCR_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return HPCR_ERR_DEVICE;
  hpcr_fuser* F = new hpcr_fuser();
  F->device = device;
  F->window_us = window_us;
  if (hipSetDevice(device) != hipSuccess) {
    delete F;
    return HPCR_ERR_DEVICE;
  }
  F->th = std::thread(fuser_loop, F);
  *out = F;
  return HPCR_OK;
}

int hpcr_fuser_destroy(hpcr_fuser* F) {
  if (!F) return HPCR_OK;
  {
    std::lock_guard<std::mutex> lk(F->mu);
    if (F->attached > 0) return HPCR_ERR_STATE;
    F->stop = true;
    F->cv_work.notify_all();
  }
  F->th.join();
  (void)hipSetDevice(F->device);
  for (auto& kv : F->plans) drop_plan(kv.second);
  delete F;
  return HPCR_OK;
}

int hpcr_fuser_set_graphs(hpcr_fuser* F, int on) {
  if (!F) return HPCR_ERR_ARG;
  std::lock_guard<std::mutex> plk(F->plan_mu);
  F->graphs = on != 0;
  if (!F->graphs) {
    for (auto& kv : F->plans) drop_plan(kv.second);
    F->plans.clear();
  }
  return HPCR_OK;
}

int hpcr_fuser_get_stats(hpcr_fuser* F, hpcr_fuser_stats* out) {
  if (!F || !out) return HPCR_ERR_ARG;
  // lock order plan_mu -> mu, as the fuser's thread takes them
  std::lock_guard<std::mutex> plk(F->plan_mu);
  std::lock_guard<std::mutex> lk(F->mu);
  *out = F->st;
  return HPCR_OK;
}

int hpcr_set_fuser(hpcr_ctx* c, hpcr_fuser* F) {
  if (!c) return HPCR_ERR_ARG;
  if (F && F->device != c->device) return set_err(c, HPCR_ERR_ARG, "fuser and context are on different devices");
  fuser_detach(c);
  if (!F) return HPCR_OK;
  HIPCHK(c, hipSetDevice(c->device));
  if (!c->ev_in) HIPCHK(c, hipEventCreateWithFlags(&c->ev_in, hipEventDisableTiming));
  if (!c->ev_out) HIPCHK(c, hipEventCreateWithFlags(&c->ev_out, hipEventDisableTiming));
  {
    std::lock_guard<std::mutex> lk(F->mu);
    F->attached++;
  }
  c->fuser = F;
  return HPCR_OK;
}

int hpcr_estimate_bytes(const hpcr_geometry* geo, int mixed, long long* bytes) {
  if (!geo || !bytes || geo->maxlevel < 0 || geo->maxlevel > 12 || geo->element_nb[0] < 1 ||
      geo->element_nb[1] < 1 || geo->element_nb[2] < 1)
    return HPCR_ERR_ARG;
  if (mixed < 0 || mixed > 2) return HPCR_ERR_ARG;
  *bytes = estimate_bytes(*geo, mixed);
  return HPCR_OK;
}

int hpcr_set_image(hpcr_ctx* c, const unsigned char* img, const hpcr_material_table* tab, double voxel_size) {
  if (!c || !img || !tab || tab->n < 2 || tab->n > HPCR_MAX_PHASES) return HPCR_ERR_ARG;
  static_assert(HPCR_MAX_PHASES == MAT_MAX, "material table size");
  (void)hipSetDevice(c->device);
  Lev& F = c->lev[c->L];
  MatTable h{};
  h.n = tab->n;
  for (int i = 0; i < tab->n; i++) {
    h.gray[i] = tab->gray[i];
    h.E[i] = tab->E[i];
    h.nu[i] = tab->nu[i];
    h.gc[i] = tab->gc[i];
  }
  unsigned char* dimg = nullptr;
  MatTable* dtab = nullptr;
  int* dbad = nullptr;
  HIPCHK(c, hipMalloc((void**)&dimg, F.N));
  hipError_t e = hipMalloc((void**)&dtab, sizeof(MatTable) + 256);
  if (e == hipSuccess) {
    dbad = (int*)((char*)dtab + sizeof(MatTable));
    e = hipMemcpyAsync(dimg, img, F.N, hipMemcpyHostToDevice, c->stream);
  }
  if (e == hipSuccess) e = hipMemcpyAsync(dtab, &h, sizeof(MatTable), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(dbad, 0, sizeof(int), c->stream);
  if (e == hipSuccess) {
    launch_image_mp(c->stream, F.N, dimg, dtab, voxel_size, F.K, F.G, F.gc, dbad);
    e = hipGetLastError();
  }
  int bad = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&bad, dbad, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(dimg);
  if (dtab) (void)hipFree(dtab);
  if (e != hipSuccess) return set_err(c, HPCR_ERR_DEVICE, std::string("set_image: ") + hipGetErrorString(e));
  if (bad) return set_err(c, HPCR_ERR_ARG, "gray level outside the material table");
  for (int f : {HPCR_K, HPCR_G, HPCR_GC}) {
    int rc = hpcr_coarse_vr(c, f);
    if (rc) return rc;
  }
  return HPCR_OK;
}

int hpcr_set_image_noalloc(hpcr_ctx* c, const unsigned char* img, const hpcr_material_table* tab,
                           double voxel_size) {
  if (!c || !img || !tab) return HPCR_ERR_ARG;
  if (tab->n < 2 || tab->n > HPCR_MAX_PHASES)
    return set_err(c, HPCR_ERR_ARG, "set_image_noalloc: the material table must have 2 to 256 rows "
                                    "(its device copy lives in a fixed scratch area)");
  if (hipSetDevice(c->device) != hipSuccess) return set_err(c, HPCR_ERR_DEVICE, "hipSetDevice");
  unsigned char* scr = nullptr;
  int rc = scratch_of(c, &scr);
  if (rc) return rc;
  Lev& F = c->lev[c->L];
  MatTable h{};
  h.n = tab->n;
  for (int i = 0; i < tab->n; i++) {
    h.gray[i] = tab->gray[i];
    h.E[i] = tab->E[i];
    h.nu[i] = tab->nu[i];
    h.gc[i] = tab->gc[i];
  }
  MatTable* dtab = (MatTable*)(scr + NOALLOC_TAB_OFF);
  int* dbad = (int*)(scr + NOALLOC_BAD_OFF);
  unsigned char* dimg = scr + NOALLOC_IMG_OFF;
  // the scratch keeps the previous call's contents: the flag starts at zero
  // (a kernel, not a memset, so that no memset node can end up in a graph)
  hipError_t e = hipMemcpyAsync(dtab, &h, sizeof(MatTable), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    zero_flag_kernel<<<dim3(1), dim3(64), 0, c->stream>>>(dbad);
    e = hipGetLastError();
  }
  // the image streams through the staging area; stream order keeps each
  // chunk's copy behind the previous chunk's kernel
  for (size_t off = 0; off < F.N && e == hipSuccess; off += NOALLOC_STAGE) {
    const size_t len = std::min(NOALLOC_STAGE, F.N - off);
    e = hipMemcpyAsync(dimg, img + off, len, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
      launch_image_mp(c->stream, len, dimg, dtab, voxel_size, F.K + off, F.G + off, F.gc + off, dbad);
      e = hipGetLastError();
    }
  }
  int bad = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&bad, dbad, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_err(c, HPCR_ERR_DEVICE, std::string("set_image_noalloc: ") + hipGetErrorString(e));
  if (bad) return set_err(c, HPCR_ERR_ARG, "gray level outside the material table");
  for (int f : {HPCR_K, HPCR_G, HPCR_GC}) {
    rc = hpcr_coarse_vr(c, f);
    if (rc) return rc;
  }
  return HPCR_OK;
}

int hpcr_u_entire(hpcr_ctx* c, double Ut, int t, double* f_norm, hpcr_stats* st) {
  if (!c || !f_norm) return HPCR_ERR_ARG;
  (void)hipSetDevice(c->device);
  if (st) memset(st, 0, sizeof(*st));
  return u_entire(c, Ut, t, f_norm, st);
}

int hpcr_u_entire_mixed(hpcr_ctx* c, double Ut, int t, double* f_norm, const hpcr_mixed_params* mp,
                        hpcr_stats* st) {
  if (!c || !f_norm || !mp) return HPCR_ERR_ARG;
  (void)hipSetDevice(c->device);
  if (st) memset(st, 0, sizeof(*st));
  return u_entire_mixed(c, Ut, t, f_norm, *mp, st);
}

int hpcr_mixed_resolve(hpcr_ctx* c, hpcr_mixed_params* mp, double* contrast) {
  if (!c || !mp) return HPCR_ERR_ARG;
  (void)hipSetDevice(c->device);
  if (contrast) *contrast = 0.0;
  return mixed_resolve(c, *mp, contrast);
}

int hpcr_phase_field(hpcr_ctx* c, int t, double* fd_norm, hpcr_stats* st) {
  if (!c || !fd_norm) return HPCR_ERR_ARG;
  (void)hipSetDevice(c->device);
  if (st) memset(st, 0, sizeof(*st));
  return phase_field(c, t, fd_norm, st);
}

int hpcr_strain_history(hpcr_ctx* c, double u_prop) {
  if (!c) return HPCR_ERR_ARG;
  (void)hipSetDevice(c->device);
  Lev& F = c->lev[c->L];
  launch_strain_history(c->stream, F.g, F.u, F.K, F.G, F.gc, F.H, c->par.lc, u_prop);
  c->diag_d_valid = false;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_launch(c, "strain_history");
}

int hpcr_voigt_all_level(hpcr_ctx* c, int f) {
  if (!c) return HPCR_ERR_ARG;
  if (f != HPCR_H && f != HPCR_KE && f != HPCR_GE) return set_err(c, HPCR_ERR_ARG, "element field expected");
  (void)hipSetDevice(c->device);
  for (int l = c->L; l > 0; l--)
    launch_elem_vr(c->stream, c->lev[l - 1].g, c->lev[l].g, field_ptr(c, f, l), field_ptr(c, f, l - 1), 1);
  c->diag_u_valid = c->diag_d_valid = false;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_launch(c, "voigt_all_level");
}

int hpcr_vr_all_level(hpcr_ctx* c, int f) {
  if (!c) return HPCR_ERR_ARG;
  if (f != HPCR_H && f != HPCR_KE && f != HPCR_GE) return set_err(c, HPCR_ERR_ARG, "element field expected");
  (void)hipSetDevice(c->device);
  for (int l = c->L; l > 0; l--)
    launch_elem_vr(c->stream, c->lev[l - 1].g, c->lev[l].g, field_ptr(c, f, l), field_ptr(c, f, l - 1), 0);
  c->diag_u_valid = c->diag_d_valid = false;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_launch(c, "vr_all_level");
}

int hpcr_fd_h(hpcr_ctx* c, int l) {
  if (!c || l < 0 || l > c->L) return HPCR_ERR_ARG;
  (void)hipSetDevice(c->device);
  launch_fd_h(c->stream, c->lev[l].g, c->lev[l].H, c->lev[l].fd);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_launch(c, "fd_h");
}

int hpcr_U_e(hpcr_ctx* c, double* Ue) {
  if (!c || !Ue) return HPCR_ERR_ARG;
  (void)hipSetDevice(c->device);
  Lev& F = c->lev[c->L];
  launch_elastic_limit(c->stream, F.g, F.u, F.K, F.G, F.gc, c->par.lc, red(c));
  int rc = read_scal(c);
  if (rc) return rc;
  *Ue = c->host_scal[S_RED];
  return check_launch(c, "U_e");
}

int hpcr_externe_force_sum(hpcr_ctx* c, double* s) {
  if (!c || !s) return HPCR_ERR_ARG;
  (void)hipSetDevice(c->device);
  Lev& F = c->lev[c->L];
  launch_top_sum(c->stream, F.g, F.fu + 2 * F.N, red(c));
  int rc = read_scal(c);
  if (rc) return rc;
  *s = c->host_scal[S_RED];
  return check_launch(c, "externe_force_sum");
}

int hpcr_delta_d_max(hpcr_ctx* c, double* out) {
  if (!c || !out) return HPCR_ERR_ARG;
  (void)hipSetDevice(c->device);
  Lev& F = c->lev[c->L];
  launch_max_diff(c->stream, F.N, F.d, F.dprev, red(c));
  int rc = read_scal(c);
  if (rc) return rc;
  *out = c->host_scal[S_RED];
  return check_launch(c, "delta_d_max");
}

int hpcr_d_max(hpcr_ctx* c, double* out) {
  if (!c || !out) return HPCR_ERR_ARG;
  (void)hipSetDevice(c->device);
  Lev& F = c->lev[c->L];
  launch_max_diff(c->stream, F.N, F.d, nullptr, red(c));
  int rc = read_scal(c);
  if (rc) return rc;
  *out = c->host_scal[S_RED];
  return check_launch(c, "d_max");
}

int hpcr_save_d_old(hpcr_ctx* c) {
  if (!c) return HPCR_ERR_ARG;
  (void)hipSetDevice(c->device);
  Lev& F = c->lev[c->L];
  HIPCHK(c, hipMemcpyAsync(F.dprev, F.d, F.N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HPCR_OK;
}

int hpcr_internal_force(hpcr_ctx* c, int l, int src, int dst) {
  if (!c || l < 0 || l > c->L || !field_vec(src) || !field_vec(dst) || src == dst) return HPCR_ERR_ARG;
  if (!field_ptr(c, src, l) || !field_ptr(c, dst, l)) return set_err(c, HPCR_ERR_ARG, "field not held on level");
  (void)hipSetDevice(c->device);
  launch_mech_apply(c->stream, c->lev[l].g, mcoef(c, l), field_ptr(c, src, l), field_ptr(c, dst, l));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_launch(c, "internal_force");
}

int hpcr_internal_force_f32(hpcr_ctx* c, int src, int dst, int mfma) {
  if (!c || !field_vec(src) || !field_vec(dst) || src == dst || (mfma != 0 && mfma != 1)) return HPCR_ERR_ARG;
  const int L = c->L;
  if (L < 1) return set_err(c, HPCR_ERR_ARG, "the FP32 preconditioner needs at least two levels");
  if (!field_ptr(c, src, L) || !field_ptr(c, dst, L)) return set_err(c, HPCR_ERR_ARG, "field not held on level");
  (void)hipSetDevice(c->device);
  int rc = mix_alloc<float>(c);
  if (rc) return rc;
  Lev& F = c->lev[L];
  MixLev<float>& M = c->mix[L];
  const size_t n = 3 * F.N;
  // the coefficients as the mixed solve's set-up forms them (split flag and
  // d as they stand), s = src in FP32, r = 0, then r -= A s (masked)
  launch_convert(c->stream, F.N, F.K, M.Kn);
  launch_convert(c->stream, F.N, F.G, M.Gn);
  launch_convert(c->stream, F.N, F.d, M.dn);
  if (mfma && (rc = mix_lm(c))) return rc;
  launch_convert(c->stream, n, (const double*)field_ptr(c, src, L), M.dA);
  HIPCHK(c, hipMemsetAsync(M.r, 0, n * sizeof(float), c->stream));
  MixCoefT<float> C = mixcoef<float>(c, L);
  C.mfma = mfma;
  launch_mix_res(c->stream, F.g, C, (const float*)M.dA, M.r);
  launch_convert(c->stream, n, (const float*)M.r, field_ptr(c, dst, L));
  c->mix_epoch++;  // the coefficients were rewritten: a mixed-form batch graph must not replay
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_launch(c, "internal_force_f32");
}

int hpcr_prem_u(hpcr_ctx* c, int l, int dst) {
  if (!c || l < 0 || l > c->L || !field_vec(dst) || !field_ptr(c, dst, l)) return HPCR_ERR_ARG;
  (void)hipSetDevice(c->device);
  int rc = mech_diag(c, l, field_ptr(c, dst, l));
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HPCR_OK;
}

int hpcr_fld(hpcr_ctx* c, int l, int src, int dst) {
  if (!c || l < 0 || l > c->L || field_vec(src) || field_vec(dst) || src == dst) return HPCR_ERR_ARG;
  if (!field_ptr(c, src, l) || !field_ptr(c, dst, l)) return set_err(c, HPCR_ERR_ARG, "field not held on level");
  (void)hipSetDevice(c->device);
  launch_pf_apply(c->stream, c->lev[l].g, pcoef(c, l), field_ptr(c, src, l), field_ptr(c, dst, l));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_launch(c, "fld");
}

int hpcr_precondition_m(hpcr_ctx* c, int l, int dst) {
  if (!c || l < 0 || l > c->L || field_vec(dst) || !field_ptr(c, dst, l)) return HPCR_ERR_ARG;
  (void)hipSetDevice(c->device);
  int rc = pf_diag(c, l, field_ptr(c, dst, l));
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HPCR_OK;
}

int hpcr_mp_choice(hpcr_ctx* c) {
  if (!c) return HPCR_ERR_ARG;
  (void)hipSetDevice(c->device);
  Lev& F = c->lev[c->L];
  // the means are recomputed as a by-product; only the flag is the result
  launch_flag_means(c->stream, F.g, F.u, F.K, F.G, F.d, F.flag, F.Ke, F.Ge, c->par.small_k, 1);
  c->diag_u_valid = false;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_launch(c, "mp_choice");
}

int hpcr_mp_all_element(hpcr_ctx* c) {
  if (!c) return HPCR_ERR_ARG;
  (void)hipSetDevice(c->device);
  Lev& F = c->lev[c->L];
  launch_flag_means(c->stream, F.g, F.u, F.K, F.G, F.d, F.flag, F.Ke, F.Ge, c->par.small_k, 0);
  c->diag_u_valid = false;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_launch(c, "mp_all_element");
}

int hpcr_mech_setup(hpcr_ctx* c) {
  if (!c) return HPCR_ERR_ARG;
  (void)hipSetDevice(c->device);
  int rc = mech_setup(c);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return HPCR_OK;
}

int hpcr_pcg_u(hpcr_ctx* c, int l, int nb, double Ut, int t, double* f_norm) {
  if (!c || l < 0 || l > c->L || nb < 0 || !f_norm) return HPCR_ERR_ARG;
  (void)hipSetDevice(c->device);
  c->diag_u_valid = false;
  int rc = pcg_u(c, l, nb, Ut, t);
  if (rc) return rc;
  if ((rc = read_scal(c))) return rc;
  if (t == 0 && l == c->L) *f_norm = c->host_scal[S_FF];
  return HPCR_OK;
}

int hpcr_pcg_d(hpcr_ctx* c, int l, int nb, double* fd_norm) {
  if (!c || l < 0 || l > c->L || nb < 0 || !fd_norm) return HPCR_ERR_ARG;
  (void)hipSetDevice(c->device);
  c->diag_d_valid = false;
  int rc = pcg_d(c, l, nb);
  if (rc) return rc;
  if ((rc = read_scal(c))) return rc;
  *fd_norm = c->host_scal[S_FF];
  return HPCR_OK;
}

int hpcr_cycle_u(hpcr_ctx* c, int l, double Ut, int t, double* f_norm) {
  if (!c || l < 0 || l > c->L || !f_norm) return HPCR_ERR_ARG;
  (void)hipSetDevice(c->device);
  int rc = mech_diag_all(c);
  if (rc) return rc;
  c->last_rr_level = -1;
  rc = cycle_u(c, l, Ut, t);
  c->diag_u_valid = false;
  if (rc) return rc;
  if ((rc = read_scal(c))) return rc;
  if (t == 0 && l == c->L) *f_norm = c->host_scal[S_FF];
  return HPCR_OK;
}

int hpcr_cycle_d(hpcr_ctx* c, int l, double* fd_norm) {
  if (!c || l < 0 || l > c->L || !fd_norm) return HPCR_ERR_ARG;
  (void)hipSetDevice(c->device);
  int rc = pf_diag_all(c);
  if (rc) return rc;
  rc = cycle_d(c, l);
  c->diag_d_valid = false;
  if (rc) return rc;
  if ((rc = read_scal(c))) return rc;
  *fd_norm = c->host_scal[S_FF];
  return HPCR_OK;
}

int hpcr_injection(hpcr_ctx* c, int lc, int f, int comp) {
  if (!c || lc < 0 || lc >= c->L) return HPCR_ERR_ARG;
  int rc = check_fl(c, f, lc, comp);
  if (rc) return rc;
  if (!field_ptr(c, f, lc + 1)) return set_err(c, HPCR_ERR_ARG, "field not held on the fine level");
  (void)hipSetDevice(c->device);
  Lev& k = c->lev[lc];
  Lev& fl = c->lev[lc + 1];
  launch_inject(c->stream, k.g, fl.g, 1, field_ptr(c, f, lc + 1) + (size_t)comp * fl.N,
                field_ptr(c, f, lc) + (size_t)comp * k.N, nullptr, 0, 0.0);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_launch(c, "injection");
}

int hpcr_restrict(hpcr_ctx* c, int lf, int rf, int fic, int dst, int comp) {
  if (!c || lf < 1 || lf > c->L) return HPCR_ERR_ARG;
  int rc;
  if ((rc = check_fl(c, rf, lf, comp)) || (rc = check_fl(c, fic, lf - 1, comp)) || (rc = check_fl(c, dst, lf - 1, comp)))
    return rc;
  (void)hipSetDevice(c->device);
  Lev& k = c->lev[lf - 1];
  Lev& f = c->lev[lf];
  launch_restrict(c->stream, k.g, f.g, 1, field_ptr(c, rf, lf) + (size_t)comp * f.N,
                  field_ptr(c, fic, lf - 1) + (size_t)comp * k.N, field_ptr(c, dst, lf - 1) + (size_t)comp * k.N);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_launch(c, "restrict");
}

int hpcr_prolong(hpcr_ctx* c, int lf, int uf, int uc, int ucold, int comp) {
  if (!c || lf < 1 || lf > c->L) return HPCR_ERR_ARG;
  int rc;
  if ((rc = check_fl(c, uf, lf, comp)) || (rc = check_fl(c, uc, lf - 1, comp)) || (rc = check_fl(c, ucold, lf - 1, comp)))
    return rc;
  (void)hipSetDevice(c->device);
  Lev& k = c->lev[lf - 1];
  Lev& f = c->lev[lf];
  double* tmp = nullptr;
  if (hipMallocAsync((void**)&tmp, k.N * sizeof(double), c->stream) != hipSuccess)
    return set_err(c, HPCR_ERR_MEMORY, "prolong scratch");
  launch_prolong(c->stream, f.g, k.g, 1, field_ptr(c, uf, lf) + (size_t)comp * f.N,
                 field_ptr(c, uc, lf - 1) + (size_t)comp * k.N, field_ptr(c, ucold, lf - 1) + (size_t)comp * k.N, 0,
                 0.0, tmp);
  (void)hipFreeAsync(tmp, c->stream);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_launch(c, "prolong");
}

int hpcr_dirichlet_bd(hpcr_ctx* c, int l, int f, double Ut) {
  if (!c || l < 0 || l > c->L || !field_vec(f) || !field_ptr(c, f, l)) return HPCR_ERR_ARG;
  (void)hipSetDevice(c->device);
  Lev& v = c->lev[l];
  launch_dirichlet_bd(c->stream, v.g, field_ptr(c, f, l), Ut, v.kplanes.data(), (int)v.kplanes.size());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_launch(c, "dirichlet_bd");
}

int hpcr_timing_enable(hpcr_ctx* c, int on) {
  if (!c) return HPCR_ERR_ARG;
  c->timing = on != 0;
  return HPCR_OK;
}
int hpcr_timing_reset(hpcr_ctx* c) {
  if (!c) return HPCR_ERR_ARG;
  (void)hipStreamSynchronize(c->stream);
  for (auto& r : c->trec) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  c->trec.clear();
  return HPCR_OK;
}
int hpcr_timing_query(hpcr_ctx* c, int kernel, int level, double* total_ms, long long* count) {
  if (!c || !total_ms || !count) return HPCR_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double tot = 0.0;
  long long n = 0;
  for (auto& r : c->trec) {
    if (r.kernel != kernel || r.level != level) continue;
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, r.a, r.b));
    tot += ms;
    n++;
  }
  *total_ms = tot;
  *count = n;
  return HPCR_OK;
}

}  // extern "C"
