// hpcrack_run — host driver of the staggered phase-field fracture loop over
// the C-ABI (include/hpcrack.h, include/hpcrack_io.h).  It restates the
// reference's main (src/main.cpp:23-107) and crack_propagation
// (src/crack_propagation.cpp:28-123): read config.txt, the uint8 image and the
// material CSV, build the level stack, then per load step solve u (unit top
// displacement), update the strain history, solve d, and write F_U.dat.
//
// usage: hpcrack_run [config.txt] [options]
//   --csv PATH          material table (default data/Material_Property.csv)
//   --steps N           maximum load steps (default 5000)
//   --lenient           accept "key = value;" lines (default: reference parsing)
//   --small-k K         residual stiffness (default 1e-5)
//   --tol-u X           squared relative residual of u (default 1e-6)
//   --tol-d X           squared relative and absolute residual of d (default 1e-12)
//   --du X              fixed top-displacement increment instead of the
//                       reference's Ue-based schedule (first load = X)
//   --du-ue F           fixed increment F*Ue (first load F*Ue), Ue from the
//                       unit-load solution as in the reference schedule
//   --stop-force F      stop when the reaction falls below F (default 0.5)
//   --out PATH          F_U file (default F_U.dat)
//   --device N          HIP device (default 0)
//   --mixed             mechanical solves in the mixed-precision form
//   --graphs            replay the multigrid cycles as HIP graphs
//   --write-history     write u, v, w, H, d every --history-every steps
//                       (default: when the config sets outTdata=1, every
//                       Tdatafrequence steps), as write_vector_3D does
//                       (src/crack_propagation.cpp:101-108), plus the loop
//                       state of that step
//   --history-every N   history interval in load steps
//   --history-dir DIR   history folder (default Time_step_data)
//   --resume T          start from the history of step T and continue at
//                       step T+1, appending to the F_U file
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "hpcrack.h"
#include "hpcrack_io.h"

namespace {

int fail(const char* what, int rc, hpcr_ctx* ctx = nullptr) {
  std::cerr << "hpcrack_run: " << what << " failed (" << rc << ")";
  const char* e = hpcr_last_error(ctx);
  if (e && *e) std::cerr << ": " << e;
  std::cerr << std::endl;
  return 1;
}

}  // namespace

int main(int argc, char** argv) {
  std::string cfg_path = "config/config.txt", csv_path = "data/Material_Property.csv", out_path = "F_U.dat";
  int steps = 5000, device = 0;
  bool strict = true;
  double small_k = 1e-5, tol_u = 1e-6, tol_d = 1e-12, du = 0.0, du_ue = 0.0, stop_force = 0.5;
  int write_history = -1, history_every = 0, resume = -1;
  bool mixed = false, graphs = false;
  std::string history_dir = "Time_step_data";
  for (int a = 1; a < argc; a++) {
    std::string s = argv[a];
    auto next = [&](const char* name) -> const char* {
      if (a + 1 >= argc) {
        std::cerr << "missing value for " << name << std::endl;
        std::exit(2);
      }
      return argv[++a];
    };
    if (s == "--csv") csv_path = next("--csv");
    else if (s == "--steps") steps = std::atoi(next("--steps"));
    else if (s == "--lenient") strict = false;
    else if (s == "--small-k") small_k = std::atof(next("--small-k"));
    else if (s == "--tol-u") tol_u = std::atof(next("--tol-u"));
    else if (s == "--tol-d") tol_d = std::atof(next("--tol-d"));
    else if (s == "--du") du = std::atof(next("--du"));
    else if (s == "--du-ue") du_ue = std::atof(next("--du-ue"));
    else if (s == "--stop-force") stop_force = std::atof(next("--stop-force"));
    else if (s == "--out") out_path = next("--out");
    else if (s == "--device") device = std::atoi(next("--device"));
    else if (s == "--mixed") mixed = true;
    else if (s == "--graphs") graphs = true;
    else if (s == "--write-history") write_history = 1;
    else if (s == "--history-every") history_every = std::atoi(next("--history-every"));
    else if (s == "--history-dir") history_dir = next("--history-dir");
    else if (s == "--resume") resume = std::atoi(next("--resume"));
    else if (s.size() > 2 && s[0] == '-' && s[1] == '-') {
      std::cerr << "unknown option " << s << std::endl;
      return 2;
    } else cfg_path = s;
  }

  hpcr_config cfg;
  int rc = hpcr_read_config(cfg_path.c_str(), strict ? 1 : 0, &cfg);
  if (rc) return fail("reading the configuration", rc);
  // one GPU holds the whole grid: the rank grid np must be 1,1,1 or absent
  // (the reference checks np against the MPI size, src/main.cpp:36-46)
  if ((cfg.seen >> HPCR_KEY_NP) & 1ULL)
    if (cfg.np[0] * std::max(cfg.np[1], 1) * std::max(cfg.np[2], 1) != 1) {
      std::cerr << "Wrong NB of MPI tasks: this solver runs one process on one GPU (np=1,1,1)" << std::endl;
      return 1;
    }
  hpcr_geometry geo;
  double lc = 0.0;
  if ((rc = hpcr_config_geometry(&cfg, &geo, &lc))) return fail("building the level stack", rc);
  int dims[3];
  for (int d = 0; d < 3; d++) dims[d] = geo.element_nb[d] * (1 << geo.maxlevel) + 1;
  const size_t N = (size_t)dims[0] * dims[1] * dims[2];
  std::cout << "ii=" << dims[0] - 1 << "\t jj=" << dims[1] - 1 << "\t kk=" << dims[2] - 1 << std::endl;

  std::vector<unsigned char> img(N);
  if ((rc = hpcr_read_raw(cfg.IMGname, cfg.voxel_nb[0], cfg.voxel_nb[1], cfg.voxel_nb[2], cfg.roi_start, dims[0],
                          dims[1], dims[2], img.data()))) {
    std::cerr << "Cannot open file!" << std::endl;
    return fail("reading the image", rc);
  }
  hpcr_material_table tab;
  if ((rc = hpcr_read_material_csv(csv_path.c_str(), &tab))) {
    std::cerr << "Could not open the file" << std::endl;
    return fail("reading the material table", rc);
  }

  hpcr_params par{};
  par.small_k = small_k;
  par.lc = lc;
  par.tol_u = tol_u;
  par.tol_d_rel = tol_d;
  par.tol_d_abs = tol_d;
  par.mg_u = cfg.mg_u;
  par.mg_d = cfg.mg_d;
  hpcr_ctx* ctx = nullptr;
  if ((rc = hpcr_ctx_create(&geo, &par, device, &ctx))) return fail("creating the device context", rc);
  if (graphs && (rc = hpcr_set_graphs(ctx, 1))) return fail("enabling graphs", rc, ctx);
  const int L = geo.maxlevel;
  // image_MP and coarse_VR on the device (input_total, src/image_input.cpp:256-282)
  if ((rc = hpcr_set_image(ctx, img.data(), &tab, cfg.voxel_size))) return fail("mapping gray levels to materials", rc, ctx);

  // history output (outTdata / Tdatafrequence, src/crack_propagation.cpp:64,101)
  if (write_history < 0) write_history = (((cfg.seen >> HPCR_KEY_OUTTDATA) & 1ULL) && cfg.outTdata == 1) ? 1 : 0;
  if (history_every <= 0)
    history_every = (((cfg.seen >> HPCR_KEY_TDATAFREQUENCE) & 1ULL) && cfg.Tdatafrequence > 0) ? cfg.Tdatafrequence : 1;
  if (write_history) std::filesystem::create_directories(history_dir);

  const double U_unit = 1.0;
  double f_norm = 0.0, fd_norm = 0.0, Ut = 0.0, delta_U = 0.0, u_prop = 1.0, Ue = 0.0;
  int t_begin = 0;
  std::vector<double> buf(N);
  static const char* const vec_names[3] = {"u", "v", "w"};
  if (resume >= 0) {
    // the fields of step `resume` and the loop state at its end
    hpcr_step_state ss;
    if ((rc = hpcr_read_step_state(history_dir.c_str(), resume, &ss))) return fail("reading the step state", rc, ctx);
    for (int c = 0; c < 3; c++) {
      if ((rc = hpcr_read_vector_3d(history_dir.c_str(), vec_names[c], resume, 0, buf.data(), N)) ||
          (rc = hpcr_set_field(ctx, HPCR_U, L, c, buf.data())))
        return fail("reading u", rc, ctx);
    }
    if ((rc = hpcr_read_vector_3d(history_dir.c_str(), "H", resume, 0, buf.data(), N)) ||
        (rc = hpcr_set_field(ctx, HPCR_H, L, 0, buf.data())))
      return fail("reading H", rc, ctx);
    if ((rc = hpcr_read_vector_3d(history_dir.c_str(), "d", resume, 0, buf.data(), N)) ||
        (rc = hpcr_set_field(ctx, HPCR_D, L, 0, buf.data())) || (rc = hpcr_save_d_old(ctx)))
      return fail("reading d", rc, ctx);
    Ut = ss.Ut;
    delta_U = ss.delta_U;
    u_prop = ss.u_prop;
    f_norm = ss.f_norm;
    Ue = ss.Ue;
    t_begin = resume + 1;
    std::cout << "resuming after step " << resume << ", Ut=" << Ut << std::endl;
  }
  std::ofstream F_U(out_path, resume >= 0 ? std::ios::app : std::ios::trunc);
  if (!F_U.is_open()) {
    std::cout << "Can not open file!" << std::endl;
    return 1;
  }
  const auto t1 = std::chrono::steady_clock::now();
  for (int t = t_begin; t < steps; t++) {
    std::cout << "***** At Time step " << t << std::endl;
    hpcr_stats su, sd;
    if (mixed) {
      const hpcr_mixed_params mp{200, tol_u, 2, 10, 6.0, 1.1, 10};
      if ((rc = hpcr_u_entire_mixed(ctx, U_unit, t, &f_norm, &mp, &su))) return fail("u_entire (mixed)", rc, ctx);
    } else if ((rc = hpcr_u_entire(ctx, U_unit, t, &f_norm, &su))) {
      return fail("u_entire", rc, ctx);
    }
    if (t == 0) {
      if (du > 0.0) {
        Ut = du;
        delta_U = du;
      } else {
        if ((rc = hpcr_U_e(ctx, &Ue))) return fail("U_e", rc, ctx);
        Ut = du_ue > 0.0 ? du_ue * Ue : Ue;
        delta_U = du_ue > 0.0 ? du_ue * Ue : 1.5e-3 * Ue;
        std::cout << "Ue=" << Ue << std::endl;
      }
      u_prop = Ut / U_unit;
    }
    if ((rc = hpcr_strain_history(ctx, u_prop))) return fail("strain_history", rc, ctx);
    if ((rc = hpcr_voigt_all_level(ctx, HPCR_H))) return fail("Voigt_all_level", rc, ctx);
    for (int l = L; l >= 0; l--)
      if ((rc = hpcr_fd_h(ctx, l))) return fail("fd_h", rc, ctx);
    if ((rc = hpcr_phase_field(ctx, t, &fd_norm, &sd))) return fail("phase_field", rc, ctx);
    double ddmax = 0.0, dmax = 0.0, FeW = 0.0;
    if ((rc = hpcr_delta_d_max(ctx, &ddmax)) || (rc = hpcr_save_d_old(ctx)) ||
        (rc = hpcr_externe_force_sum(ctx, &FeW)) || (rc = hpcr_d_max(ctx, &dmax)))
      return fail("step reductions", rc, ctx);
    FeW = cfg.voxel_size * FeW;
    if (write_history && t % history_every == 0) {
      for (int c = 0; c < 3; c++)
        if ((rc = hpcr_get_field(ctx, HPCR_U, L, c, buf.data())) ||
            (rc = hpcr_write_vector_3d(history_dir.c_str(), vec_names[c], t, 0, buf.data(), N)))
          return fail("writing u", rc, ctx);
      if ((rc = hpcr_get_field(ctx, HPCR_H, L, 0, buf.data())) ||
          (rc = hpcr_write_vector_3d(history_dir.c_str(), "H", t, 0, buf.data(), N)) ||
          (rc = hpcr_get_field(ctx, HPCR_D, L, 0, buf.data())) ||
          (rc = hpcr_write_vector_3d(history_dir.c_str(), "d", t, 0, buf.data(), N)))
        return fail("writing H, d", rc, ctx);
    }
    Ut = Ut + delta_U;
    u_prop = Ut / U_unit;
    if (write_history && t % history_every == 0) {
      const hpcr_step_state ss{t, Ut, delta_U, u_prop, f_norm, Ue};
      if ((rc = hpcr_write_step_state(history_dir.c_str(), &ss))) return fail("writing the step state", rc, ctx);
    }
    F_U << "T= " << t << ",U= " << std::setprecision(10) << Ut << " ,Fw= " << std::setprecision(10) << u_prop * FeW
        << " ,deltaU= " << delta_U << " ,delatad_max= " << ddmax << std::endl;
    std::cout << " u cycles " << su.cycles << ", d cycles " << sd.cycles << ", dmax " << dmax << std::endl;
    if (u_prop * FeW < stop_force) break;
  }
  F_U.close();
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
  std::cout << "I am processor 0, Time = " << secs << std::endl;
  hpcr_ctx_destroy(ctx);
  return 0;
}
