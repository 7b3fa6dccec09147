// Test infrastructure only.  A C-ABI shim over the reference's MPI-free
// pointwise routines, which oracle/Makefile compiles from the reference's own
// sources where they lie (shape_function.cpp, mechanical_strain.cpp,
// mechanical_stress.cpp, grid2d.cpp).  The rest of the reference's solve path
// includes mpi.h through mpi_struct.h and is not built here.
#include <vector>
#include "shape_function.h"
#include "mechanical_strain.h"
#include "mechanical_stress.h"

extern "C" {

// Dshape_function (shape_function.cpp:16-75): N[g][a] and DN[c][g][a]
void ref_dshape_function(double hx, double hy, double hz, double* N64, double* DN192) {
  grid2d N(8, 8);
  std::vector<size_t> s = {3, 8, 8};
  grid<double> DN(s);
  Dshape_function(N, DN, hx, hy, hz);
  for (int g = 0; g < 8; g++)
    for (int a = 0; a < 8; a++) N64[g * 8 + a] = N(g, a);
  for (int c = 0; c < 3; c++)
    for (int g = 0; g < 8; g++)
      for (int a = 0; a < 8; a++) DN192[(c * 8 + g) * 8 + a] = DN(c, g, a);
}

// epsilon_routine + trace_epsilon + epsilon_dev + Sigma_g on one element:
// u24[c*8 + a] is component c at element node a (reference node order).
// Outputs per Gauss point g: eps72[g*9 + 3c + m], sig72[g*9 + 3c + m],
// tr8[g], dev8[g].
void ref_element_strain_stress(const double* u24, double hx, double hy, double hz, double K,
                               double G, double* eps72, double* sig72, double* tr8, double* dev8) {
  static const int Cx[8] = {0, 1, 1, 0, 0, 1, 1, 0};
  static const int Cy[8] = {0, 0, 1, 1, 0, 0, 1, 1};
  static const int Cz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
  grid2d N(8, 8);
  std::vector<size_t> s = {3, 8, 8};
  grid<double> DN(s);
  Dshape_function(N, DN, hx, hy, hz);
  std::vector<size_t> dims = {2, 2, 2};
  gdvec<double> u(3, dims);
  for (int c = 0; c < 3; c++)
    for (int a = 0; a < 8; a++) u.getgrid(c)(Cx[a], Cy[a], Cz[a]) = u24[c * 8 + a];
  for (int g = 0; g < 8; g++) {
    grid2d eps(3, 3), sig(3, 3);
    epsilon_routine(eps, u, DN, 0, 0, 0, g);
    double tr = trace_epsilon(eps);
    Sigma_g(sig, eps, K, G, tr);
    tr8[g] = tr;
    dev8[g] = epsilon_dev(eps);
    for (int c = 0; c < 3; c++)
      for (int m = 0; m < 3; m++) {
        eps72[g * 9 + 3 * c + m] = eps(c, m);
        sig72[g * 9 + 3 * c + m] = sig(c, m);
      }
  }
}

}  // extern "C"
