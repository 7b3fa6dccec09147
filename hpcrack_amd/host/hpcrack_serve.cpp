// hpcrack_serve — request server for independent single-increment fracture
// solves on one GPU, over the C-ABI (include/hpcrack.h, include/hpcrack_io.h).
//
// A request is one seeded two-phase voxel image (node dims per axis from a
// size set, inclusion volume fraction uniform in a range) and one load step
// of crack_propagation (src/crack_propagation.cpp:65-121): the unit-load
// mechanical solve, Ue, the strain history at u_prop = load_factor * Ue, the
// phase-field solve and the reaction Fw.  Requests arrive as a Poisson
// stream.  Continuous batching: `slots` worker threads each take the oldest
// waiting request as soon as they are free and the HBM admission budget
// holds its device footprint (hpcr_estimate_bytes); every admitted request
// owns a context (its level hierarchy and strain history stay resident in
// HBM for the request's lifetime) and a HIP stream of its own, so the work
// of all admitted requests overlaps on the device.  Results do not depend on
// what else runs (every kernel is deterministic and no state is shared).
//
// Latency is measured from the request's arrival to its completion.  Images
// are generated before the clock starts (they are the request payload).
//
// usage: hpcrack_serve [options]
//   --requests N        number of requests (default 256)
//   --rate R            Poisson arrival rate per second (default 20)
//   --seed S            seed of the request stream (default 3000)
//   --sizes a,b,..      node counts per axis to draw from (default 33,65,129,257)
//   --vf lo,hi          inclusion volume fraction range (default 0.10,0.30)
//   --slots K           concurrent requests (default 8)
//   --budget-gb B       HBM admission budget (default 200)
//   --mixed             mechanical solve in the mixed-precision form (default:
//                       it reaches the tolerance where the FAS cycle stops at
//                       its cycle cap)
//   --fp64              mechanical solve with the reference's FAS cycle
//   --cheb D[,R[,U]]    mixed form: Chebyshev degree D, smoothing range R and
//                       safety factor U on the largest eigenvalue (default 0:
//                       chosen per request from its modulus contrast)
//   --coarse-iters C    mixed form: iterations of the preconditioner's
//                       coarsest solve (default: nu0_u = 50)
//   --no-graphs         launch the multigrid cycles kernel by kernel instead of
//                       replaying them as HIP graphs
//   --fuse [WINDOW_US]  attach every admitted request to one fusion engine
//                       (hpcr_fuser): the cycles of requests with the same
//                       level structure and size class, submitted within
//                       WINDOW_US (default 200) of each other, run as
//                       multi-problem launches (same-level sweeps of several
//                       requests in one launch), each batch replayed as a HIP
//                       graph unless --no-graphs; a request with no such
//                       partner runs its cycles on its own stream
//   --load-factor F     u_prop = F * Ue (default 2)
//   --tol X             squared relative residual of both solves (default 1e-16)
//   --only ID           run request ID alone (same inputs), nothing else
//   --out PATH          per-request results (default: none)
//   --image-threads T   host threads that build the request images before the
//                       clock starts (default: up to 16)
//   --hw-queues Q       hardware queues the HIP runtime spreads the slots'
//                       streams over (GPU_MAX_HW_QUEUES, 1..32; default: the
//                       runtime's own, 4)
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "hpcrack.h"
#include "hpcrack_io.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  int id = 0;
  double arrival = 0.0;  // seconds after the start of the stream
  int dims[3] = {0, 0, 0};
  hpcr_geometry geo{};
  double vf_target = 0.0, vf = 0.0;
  double rmin = 2.0, rmax = 2.0;
  unsigned long long seed = 0;
  std::vector<unsigned char> img;
  long long bytes = 0;
};

struct Result {
  int id = 0;
  int rc = 0;
  double admit = 0.0, done = 0.0;  // seconds after the start of the stream
  double Ue = 0.0, Fw = 0.0, dmax = 0.0, f_norm = 0.0, fd_norm = 0.0;
  int cyc_u = 0, cyc_d = 0;
  double rel_u = 0.0;
};

struct Options {
  int requests = 256;
  double rate = 20.0;
  unsigned long long seed = 3000;
  std::vector<int> sizes = {33, 65, 129, 257};
  double vf_lo = 0.10, vf_hi = 0.30;
  int slots = 8;
  double budget_gb = 200.0;
  bool mixed = true;
  int cheb_degree = 0;
  double cheb_ratio = 6.0, cheb_upper = 1.1;
  int image_threads = 0;
  int hw_queues = 0;
  int coarse_iters = 0;
  bool graphs = true;
  int fuse_window_us = -1;  // < 0: no fusion
  double load_factor = 2.0;
  double tol = 1e-16;
  int only = -1;
  std::string out;
};

std::vector<int> parse_ints(const char* s) {
  std::vector<int> v;
  for (const char* p = s; *p;) {
    v.push_back(std::atoi(p));
    const char* c = std::strchr(p, ',');
    if (!c) break;
    p = c + 1;
  }
  return v;
}

// levels: the coarsest level has at least 4 elements on the shortest axis
// (element_nb = (n - 1) / 2^maxlevel, e.g. 65x65x129 -> 4,4,8)
hpcr_geometry geometry_for(const int dims[3]) {
  hpcr_geometry g{};
  const int nmin = std::min(dims[0], std::min(dims[1], dims[2]));
  int m = 0;
  while (((nmin - 1) >> (m + 1)) >= 4 && ((nmin - 1) % (1 << (m + 1))) == 0) m++;
  g.maxlevel = m;
  for (int d = 0; d < 3; d++) g.element_nb[d] = (dims[d] - 1) >> m;
  g.xa = 0.0;
  g.xb = dims[0] - 1;
  g.ya = 0.0;
  g.yb = dims[1] - 1;
  g.za = 0.0;
  g.zb = dims[2] - 1;
  return g;
}

hpcr_params params_for(const Options& o) {
  hpcr_params p{};
  p.small_k = 1e-6;
  p.lc = 2.0;
  p.tol_u = o.tol;
  p.tol_d_rel = o.tol;
  p.tol_d_abs = 1e-30;
  p.mg_u = hpcr_mg{50, 4, 2, 1, 30};
  p.mg_d = hpcr_mg{10, 2, 1, 1, 30};
  return p;
}

hpcr_material_table two_phase_table() {
  hpcr_material_table t{};
  t.n = 2;
  t.gray[0] = 0;
  t.gray[1] = 1;
  t.E[0] = 3e9;
  t.E[1] = 70e9;
  t.nu[0] = 0.3;
  t.nu[1] = 0.2;
  t.gc[0] = 100.0;
  t.gc[1] = 1000.0;
  return t;
}

std::vector<Request> make_requests(const Options& o) {
  std::mt19937_64 rng(o.seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  std::vector<Request> rs(o.requests);
  double t = 0.0;
  for (int i = 0; i < o.requests; i++) {
    Request& r = rs[i];
    r.id = i;
    t += -std::log(1.0 - U(rng)) / o.rate;
    r.arrival = t;
    for (int d = 0; d < 3; d++) r.dims[d] = o.sizes[(size_t)(U(rng) * o.sizes.size()) % o.sizes.size()];
    r.vf_target = o.vf_lo + (o.vf_hi - o.vf_lo) * U(rng);
    const int nmin = std::min(r.dims[0], std::min(r.dims[1], r.dims[2]));
    r.rmin = 2.0;
    r.rmax = std::max(2.5, nmin / 12.0);
    r.seed = rng();
    r.geo = geometry_for(r.dims);
  }
  return rs;
}

int build_image(Request& r) {
  hpcr_synth_spec s{};
  for (int d = 0; d < 3; d++) s.dims[d] = r.dims[d];
  s.seed = r.seed;
  s.matrix_label = 0;
  s.n_incl = 0;
  // spheres of label 1 placed until the volume fraction is reached
  s.pore_fraction = r.vf_target;
  s.pore_rmin = r.rmin;
  s.pore_rmax = r.rmax;
  s.pore_label = 1;
  s.gap = 1;
  s.max_tries = 4000000;
  r.img.resize((size_t)r.dims[0] * r.dims[1] * r.dims[2]);
  double frac[2] = {0, 0};
  const int rc = hpcr_synth_image(&s, r.img.data(), frac);
  r.vf = frac[1];
  return rc;
}

// one load step of crack_propagation on a fresh context
Result solve(const Request& r, const Options& o, hpcr_fuser* fz, Clock::time_point start) {
  Result res;
  res.id = r.id;
  const hpcr_params par = params_for(o);
  const hpcr_material_table tab = two_phase_table();
  hpcr_ctx* ctx = nullptr;
  int rc = hpcr_ctx_create(&r.geo, &par, 0, &ctx);
  const int L = r.geo.maxlevel;
  double f_norm = 0.0, fd_norm = 0.0, FeW = 0.0, Ue = 0.0;
  hpcr_stats su{}, sd{};
  if (!rc && fz) rc = hpcr_set_fuser(ctx, fz);
  if (!rc && o.graphs) rc = hpcr_set_graphs(ctx, 1);
  if (!rc) rc = hpcr_set_image(ctx, r.img.data(), &tab, 1e-6);
  if (!rc) {
    if (o.mixed) {
      hpcr_mixed_params mp{200, o.tol, o.cheb_degree, 10, o.cheb_ratio, o.cheb_upper, 10, 0, 0, o.coarse_iters};
      rc = hpcr_u_entire_mixed(ctx, 1.0, 0, &f_norm, &mp, &su);
    } else {
      rc = hpcr_u_entire(ctx, 1.0, 0, &f_norm, &su);
    }
  }
  if (!rc) rc = hpcr_U_e(ctx, &Ue);
  const double u_prop = o.load_factor * Ue;
  if (!rc) rc = hpcr_strain_history(ctx, u_prop);
  if (!rc) rc = hpcr_voigt_all_level(ctx, HPCR_H);
  for (int l = L; l >= 0 && !rc; l--) rc = hpcr_fd_h(ctx, l);
  if (!rc) rc = hpcr_phase_field(ctx, 0, &fd_norm, &sd);
  if (!rc) rc = hpcr_externe_force_sum(ctx, &FeW);
  if (!rc) rc = hpcr_d_max(ctx, &res.dmax);
  if (rc && ctx) std::fprintf(stderr, "request %d: %s\n", r.id, hpcr_last_error(ctx));
  if (rc && !ctx) std::fprintf(stderr, "request %d: %s\n", r.id, hpcr_last_error(nullptr));
  if (ctx) hpcr_ctx_destroy(ctx);
  res.rc = rc;
  res.Ue = Ue;
  res.Fw = u_prop * 1e-6 * FeW;
  res.f_norm = f_norm;
  res.fd_norm = fd_norm;
  res.cyc_u = su.cycles;
  res.cyc_d = sd.cycles;
  res.rel_u = f_norm != 0.0 ? su.r_norm_final / f_norm : 0.0;
  res.done = std::chrono::duration<double>(Clock::now() - start).count();
  return res;
}

double percentile(std::vector<double> v, double p) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double k = p * (v.size() - 1);
  const size_t i = (size_t)std::floor(k);
  const size_t j = std::min(i + 1, v.size() - 1);
  return v[i] + (k - i) * (v[j] - v[i]);
}

}  // namespace

int main(int argc, char** argv) {
  Options o;
  for (int a = 1; a < argc; a++) {
    std::string s = argv[a];
    auto next = [&]() -> const char* {
      if (a + 1 >= argc) {
        std::fprintf(stderr, "missing value for %s\n", s.c_str());
        std::exit(2);
      }
      return argv[++a];
    };
    if (s == "--requests") o.requests = std::atoi(next());
    else if (s == "--rate") o.rate = std::atof(next());
    else if (s == "--seed") o.seed = std::strtoull(next(), nullptr, 10);
    else if (s == "--sizes") o.sizes = parse_ints(next());
    else if (s == "--vf") {
      const char* v = next();
      o.vf_lo = std::atof(v);
      const char* c = std::strchr(v, ',');
      o.vf_hi = c ? std::atof(c + 1) : o.vf_lo;
    } else if (s == "--slots") o.slots = std::atoi(next());
    else if (s == "--budget-gb") o.budget_gb = std::atof(next());
    else if (s == "--mixed") o.mixed = true;
    else if (s == "--fp64") o.mixed = false;
    else if (s == "--cheb") {
      const char* v = next();
      o.cheb_degree = std::atoi(v);
      const char* c = std::strchr(v, ',');
      if (c) {
        o.cheb_ratio = std::atof(c + 1);
        c = std::strchr(c + 1, ',');
        if (c) o.cheb_upper = std::atof(c + 1);
      }
    } else if (s == "--image-threads") o.image_threads = std::atoi(next());
    else if (s == "--hw-queues") o.hw_queues = std::atoi(next());
    else if (s == "--coarse-iters") o.coarse_iters = std::atoi(next());
    else if (s == "--no-graphs") o.graphs = false;
    else if (s == "--fuse") {
      o.fuse_window_us = 200;
      if (a + 1 < argc && argv[a + 1][0] != '-') o.fuse_window_us = std::atoi(argv[++a]);
    }
    else if (s == "--load-factor") o.load_factor = std::atof(next());
    else if (s == "--tol") o.tol = std::atof(next());
    else if (s == "--only") o.only = std::atoi(next());
    else if (s == "--out") o.out = next();
    else {
      std::fprintf(stderr, "unknown option %s\n", s.c_str());
      return 2;
    }
  }
  if (o.requests < 1 || o.rate <= 0 || o.sizes.empty() || o.slots < 1 || o.budget_gb <= 0 || o.cheb_degree < 0 ||
      o.cheb_degree > 8 || !(o.cheb_ratio > 1.0) || !(o.cheb_upper > 0.0) || o.coarse_iters < 0) {
    std::fprintf(stderr, "bad options\n");
    return 2;
  }
  if (o.hw_queues < 0 || o.hw_queues > 32) {
    std::fprintf(stderr, "--hw-queues takes 1..32\n");
    return 2;
  }
  if (o.hw_queues > 0) {
    // read by the HIP runtime when it initialises (the first HIP call below)
    const std::string q = std::to_string(o.hw_queues);
    setenv("GPU_MAX_HW_QUEUES", q.c_str(), 1);
  }
  for (int n : o.sizes)
    if (n < 9 || ((n - 1) & 7) != 0) {
      std::fprintf(stderr, "node counts must be 8k+1 >= 9\n");
      return 2;
    }
  int ndev = 0;
  if (hpcr_device_count(&ndev) != HPCR_OK || ndev < 1) {
    std::fprintf(stderr, "no HIP device\n");
    return 1;
  }

  std::vector<Request> reqs = make_requests(o);
  if (o.only >= 0) {
    if (o.only >= o.requests) return 2;
    reqs = {reqs[o.only]};
  }
  {
    // the request images (the payload), built on host threads before the
    // clock starts; each depends only on its own seed
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nt = o.image_threads > 0 ? o.image_threads : (int)std::min(16u, hw);
    std::atomic<size_t> next_img{0}, ready{0};
    std::atomic<int> bad{-1};
    std::vector<std::thread> builders;
    for (int t = 0; t < nt; t++)
      builders.emplace_back([&] {
        for (size_t i; (i = next_img.fetch_add(1)) < reqs.size();) {
          Request& r = reqs[i];
          if (build_image(r) || hpcr_estimate_bytes(&r.geo, o.mixed ? 1 : 0, &r.bytes)) {
            bad = r.id;
            continue;
          }
          const size_t k = ++ready;
          if (k % 32 == 0 || k == reqs.size()) {
            std::fprintf(stderr, "[serve] %zu/%zu request images ready\n", k, reqs.size());
            std::fflush(stderr);
          }
        }
      });
    for (auto& b : builders) b.join();
    if (bad >= 0) {
      std::fprintf(stderr, "image of request %d failed\n", (int)bad);
      return 1;
    }
  }
  const long long budget = (long long)(o.budget_gb * 1e9);
  for (const Request& r : reqs)
    if (r.bytes > budget) {
      std::fprintf(stderr, "request %d needs %lld bytes, above the budget\n", r.id, r.bytes);
      return 1;
    }

  hpcr_fuser* fz = nullptr;
  if (o.fuse_window_us >= 0) {
    if (hpcr_fuser_create(0, o.fuse_window_us, &fz) != HPCR_OK || hpcr_fuser_set_graphs(fz, o.graphs ? 1 : 0)) {
      std::fprintf(stderr, "cannot create the fusion engine\n");
      return 1;
    }
  }

  // ---- continuous batching ----
  std::mutex mu;
  std::condition_variable cv;
  std::deque<int> waiting;  // indices into reqs, arrival order
  long long in_use = 0, peak_bytes = 0;
  int active = 0, peak_active = 0;
  // requests that reached the front of the queue with a slot free but had
  // to wait for memory (admission held back by the budget)
  std::vector<char> held(reqs.size(), 0);
  int budget_waits = 0;
  size_t arrived = 0, completed = 0;
  std::vector<Result> results(reqs.size());
  const bool alone = o.only >= 0;
  const auto start = Clock::now();

  std::thread arrivals([&] {
    for (size_t i = 0; i < reqs.size(); i++) {
      if (!alone) std::this_thread::sleep_until(start + std::chrono::duration<double>(reqs[i].arrival));
      std::lock_guard<std::mutex> lk(mu);
      waiting.push_back((int)i);
      arrived++;
      cv.notify_all();
    }
  });
  std::vector<std::thread> workers;
  const int nslots = alone ? 1 : o.slots;
  for (int w = 0; w < nslots; w++) {
    workers.emplace_back([&] {
      for (;;) {
        int i = -1;
        {
          std::unique_lock<std::mutex> lk(mu);
          // the oldest waiting request, once the budget holds it
          cv.wait(lk, [&] {
            if (waiting.empty()) return arrived == reqs.size();
            const int f = waiting.front();
            if (in_use + reqs[f].bytes <= budget) return true;
            if (!held[f]) {
              held[f] = 1;
              budget_waits++;
            }
            return false;
          });
          if (waiting.empty()) return;
          i = waiting.front();
          waiting.pop_front();
          in_use += reqs[i].bytes;
          peak_bytes = std::max(peak_bytes, in_use);
          active++;
          peak_active = std::max(peak_active, active);
        }
        const double admit = std::chrono::duration<double>(Clock::now() - start).count();
        Result r = solve(reqs[i], o, fz, start);
        r.admit = admit;
        {
          std::lock_guard<std::mutex> lk(mu);
          in_use -= reqs[i].bytes;
          active--;
          results[i] = r;
          if (++completed % 32 == 0 || completed == reqs.size()) {
            std::fprintf(stderr, "[serve] %zu/%zu requests done at %.2f s\n", completed, reqs.size(), r.done);
            std::fflush(stderr);
          }
        }
        cv.notify_all();
      }
    });
  }
  arrivals.join();
  for (auto& t : workers) t.join();
  hpcr_fuser_stats fst{};
  if (fz) {
    hpcr_fuser_get_stats(fz, &fst);
    hpcr_fuser_destroy(fz);
  }

  // ---- report ----
  std::vector<double> lat, wait, service;
  int failed = 0, cu = 0, cd = 0;
  double last = 0.0;
  for (size_t i = 0; i < reqs.size(); i++) {
    const Result& r = results[i];
    const double arr = alone ? 0.0 : reqs[i].arrival;
    if (r.rc) failed++;
    lat.push_back(r.done - arr);
    wait.push_back(r.admit - arr);
    service.push_back(r.done - r.admit);
    last = std::max(last, r.done);
    cu += r.cyc_u;
    cd += r.cyc_d;
  }
  if (!o.out.empty()) {
    FILE* f = std::fopen(o.out.c_str(), "w");
    if (!f) return 1;
    for (size_t i = 0; i < reqs.size(); i++) {
      const Result& r = results[i];
      const Request& q = reqs[i];
      std::fprintf(f,
                   "id %d dims %dx%dx%d vf %.4f rc %d cyc_u %d cyc_d %d rel_u %.17g Ue %.17g Fw %.17g dmax %.17g "
                   "f_norm %.17g fd_norm %.17g latency_ms %.3f\n",
                   q.id, q.dims[0], q.dims[1], q.dims[2], q.vf, r.rc, r.cyc_u, r.cyc_d, r.rel_u, r.Ue, r.Fw, r.dmax,
                   r.f_norm, r.fd_norm, 1e3 * lat[i]);
    }
    std::fclose(f);
  }
  const double first = alone ? 0.0 : reqs.front().arrival;
  std::printf(
      "{\"requests\": %zu, \"failed\": %d, \"form\": \"%s\", \"slots\": %d, \"budget_gb\": %.1f, \"rate_per_s\": %.2f, "
      "\"solves_per_s\": %.3f, \"latency_ms_p50\": %.1f, \"latency_ms_p99\": %.1f, \"latency_ms_max\": %.1f, "
      "\"wait_ms_p50\": %.1f, \"service_ms_p50\": %.1f, \"peak_admitted_gb\": %.2f, \"peak_concurrent\": %d, "
      "\"mean_cycles_u\": %.2f, \"mean_cycles_d\": %.2f, \"makespan_s\": %.3f, \"fused\": %s, "
      "\"fused_batches\": %lld, \"fused_cycles\": %lld, \"fused_launches\": %lld, \"fused_problem_launches\": %lld, "
      "\"fused_max_problems\": %d, \"fused_graph_replays\": %lld, \"budget_waits\": %d, \"cheb_degree\": %d, "
      "\"hw_queues\": %d}\n",
      reqs.size(), failed, o.mixed ? "mixed" : "fp64", nslots, o.budget_gb, o.rate,
      reqs.size() / std::max(1e-9, last - first), 1e3 * percentile(lat, 0.5), 1e3 * percentile(lat, 0.99),
      1e3 * percentile(lat, 1.0), 1e3 * percentile(wait, 0.5), 1e3 * percentile(service, 0.5), peak_bytes / 1e9,
      peak_active, (double)cu / reqs.size(), (double)cd / reqs.size(), last, fz ? "true" : "false", fst.batches,
      fst.cycles, fst.launches, fst.problem_launches, fst.max_problems, fst.graph_replays, budget_waits,
      o.mixed ? o.cheb_degree : 0, o.hw_queues);
  return failed ? 1 : 0;
}
