// Host-side input formats (include/hpcrack_io.h): config.txt, material CSV,
// uint8 .raw images, the level geometry they imply, and the seeded synthetic
// image generator.  Plain C++, no device calls.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/hpcrack_io.h"

namespace {

std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

// stoi / stod that report failure instead of throwing
bool to_int(const std::string& v, int* out) {
  try {
    *out = std::stoi(v);
    return true;
  } catch (...) {
    return false;
  }
}
bool to_dbl(const std::string& v, double* out) {
  try {
    *out = std::stod(v);
    return true;
  } catch (...) {
    return false;
  }
}
template <class T, class F>
bool parse_vec(const std::string& v, T* out, int n, F conv) {
  std::stringstream ss(v);
  std::string tok;
  int i = 0;
  while (std::getline(ss, tok, ',')) {
    if (i < n && !conv(tok, &out[i])) return false;
    i++;
  }
  return i >= 1;
}

// xoshiro256** (Blackman & Vigna) seeded through splitmix64
struct Rng {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto& v : s) v = splitmix(x);
  }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
  int index(int n) {
    int v = (int)(uniform() * n);
    return v >= n ? n - 1 : v;
  }
};

struct Sphere {
  int cx, cy, cz;
  double r;
};

}  // namespace

// file name of a history field (write_read_history.cpp:18-53): NAME_t_T_Proc_P.dat
// under DIR, or NAME itself when it already names a .dat file
static std::string history_path(const char* dir, const char* name, int t, int proc) {
  if (std::strstr(name, ".dat")) return name;
  return std::string(dir ? dir : "Time_step_data") + "/" + name + "_t_" + std::to_string(t) + "_Proc_" +
         std::to_string(proc) + ".dat";
}

extern "C" {

int hpcr_read_config(const char* path, int strict, hpcr_config* out) {
  if (!path || !out) return HPCR_ERR_ARG;
  memset(out, 0, sizeof(*out));
  std::ifstream file(path);
  if (!file.is_open()) return HPCR_ERR_IO;
  std::string line;
  static const char* keys[HPCR_KEY_COUNT] = {
      "D", "np", "X_start", "IMGname", "voxel_nb_X", "voxel_nb_Y", "voxel_nb_Z", "ROI_start_X", "ROI_start_Y",
      "ROI_start_Z", "voxel_size", "element_nb", "ncy_u", "nu0_u", "nu1_u", "nu2_u", "gamma_u", "ncy_d",
      "nu0_d", "nu1_d", "nu2_d", "gamma_d", "lc", "outK", "outG", "outgc", "outdesplacement", "outH",
      "outTdata", "Tdatafrequence"};
  while (std::getline(file, line)) {
    if (line.empty() || line[0] == '#') continue;
    const size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    std::string key = line.substr(0, eq);
    std::string val = line.substr(eq + 1);
    if (!strict) {
      key = trim(key);
      val = trim(val);
      while (!val.empty() && val.back() == ';') val.pop_back();
      val = trim(val);
    }
    int k = -1;
    for (int q = 0; q < HPCR_KEY_COUNT; q++)
      if (key == keys[q]) k = q;
    if (k < 0) continue;
    bool ok = true;
    double dv = 0;
    int iv = 0;
    switch (k) {
      case HPCR_KEY_D: ok = to_int(val, &out->D); break;
      case HPCR_KEY_NP: ok = parse_vec(val, out->np, 3, to_int); break;
      case HPCR_KEY_X_START: ok = parse_vec(val, out->X_start, 3, to_dbl); break;
      case HPCR_KEY_IMGNAME: snprintf(out->IMGname, sizeof(out->IMGname), "%s", val.c_str()); break;
      case HPCR_KEY_VOXEL_NB_X: ok = to_dbl(val, &dv); out->voxel_nb[0] = (int)dv; break;
      case HPCR_KEY_VOXEL_NB_Y: ok = to_dbl(val, &dv); out->voxel_nb[1] = (int)dv; break;
      case HPCR_KEY_VOXEL_NB_Z: ok = to_dbl(val, &dv); out->voxel_nb[2] = (int)dv; break;
      case HPCR_KEY_ROI_X: ok = to_dbl(val, &dv); out->roi_start[0] = (int)dv; break;
      case HPCR_KEY_ROI_Y: ok = to_dbl(val, &dv); out->roi_start[1] = (int)dv; break;
      case HPCR_KEY_ROI_Z: ok = to_dbl(val, &dv); out->roi_start[2] = (int)dv; break;
      case HPCR_KEY_VOXEL_SIZE: ok = to_dbl(val, &out->voxel_size); break;
      case HPCR_KEY_ELEMENT_NB: ok = parse_vec(val, out->element_nb, 3, to_int); break;
      case HPCR_KEY_NCY_U: ok = to_int(val, &out->mg_u.ncy); break;
      case HPCR_KEY_NU0_U: ok = to_int(val, &out->mg_u.nu0); break;
      case HPCR_KEY_NU1_U: ok = to_int(val, &out->mg_u.nu1); break;
      case HPCR_KEY_NU2_U: ok = to_int(val, &out->mg_u.nu2); break;
      case HPCR_KEY_GAMMA_U: ok = to_int(val, &out->mg_u.gamma); break;
      case HPCR_KEY_NCY_D: ok = to_int(val, &out->mg_d.ncy); break;
      case HPCR_KEY_NU0_D: ok = to_int(val, &out->mg_d.nu0); break;
      case HPCR_KEY_NU1_D: ok = to_int(val, &out->mg_d.nu1); break;
      case HPCR_KEY_NU2_D: ok = to_int(val, &out->mg_d.nu2); break;
      case HPCR_KEY_GAMMA_D: ok = to_int(val, &out->mg_d.gamma); break;
      case HPCR_KEY_LC:
        // the reference stores stoi(value) in a double (configuration.cpp:72-73)
        if (strict) {
          ok = to_int(val, &iv);
          out->lc = iv;
        } else {
          ok = to_dbl(val, &out->lc);
        }
        break;
      case HPCR_KEY_OUTK: ok = to_int(val, &out->outK); break;
      case HPCR_KEY_OUTG: ok = to_int(val, &out->outG); break;
      case HPCR_KEY_OUTGC: ok = to_int(val, &out->outgc); break;
      case HPCR_KEY_OUTDESPLACEMENT: ok = to_int(val, &out->outdesplacement); break;
      case HPCR_KEY_OUTH: ok = to_int(val, &out->outH); break;
      case HPCR_KEY_OUTTDATA: ok = to_int(val, &out->outTdata); break;
      case HPCR_KEY_TDATAFREQUENCE: ok = to_int(val, &out->Tdatafrequence); break;
      default: break;
    }
    if (!ok) return HPCR_ERR_IO;  // the reference's stoi/stod would throw
    out->seen |= 1ULL << k;
  }
  return HPCR_OK;
}

int hpcr_config_geometry(const hpcr_config* cfg, hpcr_geometry* geo, double* lc_scaled) {
  if (!cfg || !geo) return HPCR_ERR_ARG;
  for (int d = 0; d < 3; d++)
    if (cfg->element_nb[d] < 1) return HPCR_ERR_ARG;
  // maxlevel from the X extent (main.cpp:52-68)
  int maxlevel = 0;
  const int nb_ele_f = cfg->voxel_nb[0] - cfg->roi_start[0] - 1;
  int nb_ele_c = cfg->element_nb[0];
  if (2 * nb_ele_c > nb_ele_f) {
    maxlevel = 0;
  } else {
    do {
      nb_ele_c = 2 * nb_ele_c;
      if (nb_ele_c <= nb_ele_f) maxlevel += 1;
    } while (nb_ele_c <= nb_ele_f);
  }
  // input_setting (image_input.cpp:21-29) and initialize (initialization.cpp:21-37)
  int coefb[3];
  for (int d = 0; d < 3; d++) coefb[d] = cfg->roi_start[d] + cfg->element_nb[d] * (1 << maxlevel);
  geo->maxlevel = maxlevel;
  for (int d = 0; d < 3; d++) geo->element_nb[d] = cfg->element_nb[d];
  geo->xa = cfg->X_start[0];
  geo->ya = cfg->X_start[1];
  geo->za = cfg->X_start[2];
  geo->xb = geo->xa + coefb[0];
  geo->yb = geo->ya + coefb[1];
  geo->zb = geo->za + coefb[2];
  if (lc_scaled) {
    double hz = (geo->zb - geo->za) / cfg->element_nb[2];
    for (int l = 0; l < maxlevel; l++) hz *= 0.5;
    *lc_scaled = cfg->lc * hz;
  }
  return HPCR_OK;
}

int hpcr_read_material_csv(const char* path, hpcr_material_table* out) {
  if (!path || !out) return HPCR_ERR_ARG;
  memset(out, 0, sizeof(*out));
  std::ifstream file(path);
  if (!file.is_open()) return HPCR_ERR_IO;
  std::vector<std::vector<std::string>> rows;
  std::string line, word;
  while (std::getline(file, line)) {
    std::vector<std::string> row;
    std::stringstream ss(line);
    while (std::getline(ss, word, ',')) row.push_back(word);
    rows.push_back(row);
  }
  if (rows.size() < 4) return HPCR_ERR_IO;
  const int n = (int)rows[0].size() - 1;
  if (n < 2 || n > HPCR_MAX_PHASES) return HPCR_ERR_IO;
  for (int r = 1; r <= 3; r++)
    if ((int)rows[r].size() - 1 < n) return HPCR_ERR_IO;
  for (int i = 0; i < n; i++) {
    if (!to_int(rows[0][i + 1], &out->gray[i])) return HPCR_ERR_IO;
    if (!to_dbl(rows[1][i + 1], &out->E[i]) || !to_dbl(rows[2][i + 1], &out->nu[i]) ||
        !to_dbl(rows[3][i + 1], &out->gc[i]))
      return HPCR_ERR_IO;
  }
  out->n = n;
  return HPCR_OK;
}

int hpcr_read_raw(const char* path, int vx, int vy, int vz, const int roi[3], int nx, int ny, int nz,
                  unsigned char* out) {
  if (!path || !out || !roi || vx < 1 || vy < 1 || vz < 1) return HPCR_ERR_ARG;
  if (roi[0] < 0 || roi[1] < 0 || roi[2] < 0 || roi[0] + nx > vx || roi[1] + ny > vy || roi[2] + nz > vz)
    return HPCR_ERR_ARG;
  std::ifstream is(path, std::ios::binary);
  if (!is) return HPCR_ERR_IO;
  is.seekg(0, is.end);
  const long long length = is.tellg();
  is.seekg(0, is.beg);
  const long long need = (long long)vx * vy * vz;
  if (length < need) return HPCR_ERR_IO;
  std::vector<unsigned char> buf((size_t)need);
  is.read(reinterpret_cast<char*>(buf.data()), need);
  if (!is) return HPCR_ERR_IO;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
        out[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k)] =
            buf[(size_t)(roi[0] + i) + (size_t)vx * ((size_t)(roi[1] + j) + (size_t)vy * (size_t)(roi[2] + k))];
  return HPCR_OK;
}

int hpcr_write_raw(const char* path, const unsigned char* img, size_t n) {
  if (!path || !img) return HPCR_ERR_ARG;
  FILE* f = fopen(path, "wb");
  if (!f) return HPCR_ERR_IO;
  size_t w = fwrite(img, 1, n, f);
  fclose(f);
  return w == n ? HPCR_OK : HPCR_ERR_IO;
}

int hpcr_write_vector_3d(const char* dir, const char* name, int t, int proc, const double* v, size_t n) {
  if (!name || !v) return HPCR_ERR_ARG;
  const std::string path = history_path(dir, name, t, proc);
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return HPCR_ERR_IO;
  const size_t w = fwrite(v, sizeof(double), n, f);
  const int cl = fclose(f);
  return (w == n && cl == 0) ? HPCR_OK : HPCR_ERR_IO;
}

int hpcr_read_vector_3d(const char* dir, const char* name, int t, int proc, double* v, size_t n) {
  if (!name || !v) return HPCR_ERR_ARG;
  const std::string path = history_path(dir, name, t, proc);
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return HPCR_ERR_IO;
  const size_t r = fread(v, sizeof(double), n, f);
  const bool at_end = (r == n) && (fgetc(f) == EOF);
  fclose(f);
  return at_end ? HPCR_OK : HPCR_ERR_IO;
}

int hpcr_write_step_state(const char* dir, const hpcr_step_state* s) {
  if (!s) return HPCR_ERR_ARG;
  const std::string path = std::string(dir ? dir : "Time_step_data") + "/state_t_" + std::to_string(s->t) + ".txt";
  FILE* f = fopen(path.c_str(), "w");
  if (!f) return HPCR_ERR_IO;
  fprintf(f, "t %d\nUt %.17g\ndelta_U %.17g\nu_prop %.17g\nf_norm %.17g\nUe %.17g\n", s->t, s->Ut, s->delta_U,
          s->u_prop, s->f_norm, s->Ue);
  return fclose(f) == 0 ? HPCR_OK : HPCR_ERR_IO;
}

int hpcr_read_step_state(const char* dir, int t, hpcr_step_state* s) {
  if (!s) return HPCR_ERR_ARG;
  const std::string path = std::string(dir ? dir : "Time_step_data") + "/state_t_" + std::to_string(t) + ".txt";
  std::ifstream in(path);
  if (!in) return HPCR_ERR_IO;
  hpcr_step_state v{};
  int got = 0;
  std::string key;
  while (in >> key) {
    if (key == "t") { in >> v.t; got |= 1; }
    else if (key == "Ut") { in >> v.Ut; got |= 2; }
    else if (key == "delta_U") { in >> v.delta_U; got |= 4; }
    else if (key == "u_prop") { in >> v.u_prop; got |= 8; }
    else if (key == "f_norm") { in >> v.f_norm; got |= 16; }
    else if (key == "Ue") { in >> v.Ue; got |= 32; }
    else return HPCR_ERR_IO;
    if (in.fail()) return HPCR_ERR_IO;
  }
  if (got != 63 || v.t != t) return HPCR_ERR_IO;
  *s = v;
  return HPCR_OK;
}

int hpcr_image_mp(const hpcr_material_table* tab, const unsigned char* img, size_t n, double voxel_size,
                  double* K, double* G, double* gc) {
  if (!tab || !img || !K || !G || !gc || tab->n < 2) return HPCR_ERR_ARG;
  for (size_t q = 0; q < n; q++) {
    const int x = img[q];
    int pos = -1;
    // findPosition (csv_reader.cpp:74-82)
    for (int i = 0; i < tab->n - 1; ++i)
      if (x >= tab->gray[i] && x <= tab->gray[i + 1]) {
        pos = i;
        break;
      }
    if (pos < 0 || tab->gray[pos + 1] == tab->gray[pos]) return HPCR_ERR_ARG;
    // integer division, as in image_input.cpp:134-137
    const int fac = (x - tab->gray[pos]) / (tab->gray[pos + 1] - tab->gray[pos]);
    double E = tab->E[pos] + fac * (tab->E[pos + 1] - tab->E[pos]);
    const double nu = tab->nu[pos] + fac * (tab->nu[pos + 1] - tab->nu[pos]);
    gc[q] = tab->gc[pos] + fac * (tab->gc[pos + 1] - tab->gc[pos]);
    E = voxel_size * E;
    K[q] = 1.0 / (3.0 * (1.0 - 2.0 * nu)) * E;
    G[q] = 1.0 / (2.0 * (1.0 + nu)) * E;
  }
  return HPCR_OK;
}

int hpcr_synth_image(const hpcr_synth_spec* sp, unsigned char* out, double realised[2]) {
  if (!sp || !out) return HPCR_ERR_ARG;
  const int nx = sp->dims[0], ny = sp->dims[1], nz = sp->dims[2];
  if (nx < 1 || ny < 1 || nz < 1) return HPCR_ERR_ARG;
  const size_t n = (size_t)nx * ny * nz;
  memset(out, sp->matrix_label, n);
  Rng rng(sp->seed);
  std::vector<Sphere> placed;
  const int max_tries = sp->max_tries > 0 ? sp->max_tries : 1000000;
  // Placed spheres are binned by centre on a grid of cells no smaller than
  // the largest possible centre distance of two overlapping spheres, so a
  // candidate is tested against the 27 cells around its centre only (the
  // same accept/reject decisions as testing every placed sphere).
  const double rmax = std::max(sp->n_incl > 0 ? sp->incl_rmax : 0.0, sp->pore_fraction > 0 ? sp->pore_rmax : 0.0);
  const int cell = std::max(1, (int)std::ceil(2.0 * rmax + sp->gap));
  const int gx = (nx + cell - 1) / cell, gy = (ny + cell - 1) / cell, gz = (nz + cell - 1) / cell;
  std::vector<std::vector<int>> bins((size_t)gx * gy * gz);
  auto bin_of = [&](int cx, int cy, int cz) { return (size_t)(cx / cell) + (size_t)gx * ((cy / cell) + (size_t)gy * (cz / cell)); };
  auto fits = [&](const Sphere& s) {
    const int bx = s.cx / cell, by = s.cy / cell, bz = s.cz / cell;
    for (int kz = std::max(0, bz - 1); kz <= std::min(gz - 1, bz + 1); kz++)
      for (int ky = std::max(0, by - 1); ky <= std::min(gy - 1, by + 1); ky++)
        for (int kx = std::max(0, bx - 1); kx <= std::min(gx - 1, bx + 1); kx++)
          for (int idx : bins[(size_t)kx + (size_t)gx * (ky + (size_t)gy * kz)]) {
            const Sphere& o = placed[idx];
            const double dx = s.cx - o.cx, dy = s.cy - o.cy, dz = s.cz - o.cz;
            const double lim = s.r + o.r + sp->gap;
            if (dx * dx + dy * dy + dz * dz < lim * lim) return false;
          }
    return true;
  };
  auto place = [&](const Sphere& s) {
    bins[bin_of(s.cx, s.cy, s.cz)].push_back((int)placed.size());
    placed.push_back(s);
  };
  auto stamp = [&](const Sphere& s, int label) {
    size_t cnt = 0;
    const int R = (int)std::ceil(s.r);
    for (int k = std::max(0, s.cz - R); k <= std::min(nz - 1, s.cz + R); k++)
      for (int j = std::max(0, s.cy - R); j <= std::min(ny - 1, s.cy + R); j++)
        for (int i = std::max(0, s.cx - R); i <= std::min(nx - 1, s.cx + R); i++) {
          const double dx = i - s.cx, dy = j - s.cy, dz = k - s.cz;
          if (dx * dx + dy * dy + dz * dz <= s.r * s.r) {
            out[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k)] = (unsigned char)label;
            cnt++;
          }
        }
    return cnt;
  };
  auto draw = [&](double rmin, double rmax) {
    Sphere s;
    s.cx = rng.index(nx);
    s.cy = rng.index(ny);
    s.cz = rng.index(nz);
    s.r = rng.uniform(rmin, rmax);
    return s;
  };
  size_t incl_vox = 0, pore_vox = 0;
  int tries = 0;
  for (int q = 0; q < sp->n_incl && tries < max_tries;) {
    Sphere s = draw(sp->incl_rmin, sp->incl_rmax);
    tries++;
    if (!fits(s)) continue;
    place(s);
    incl_vox += stamp(s, sp->incl_label);
    q++;
  }
  tries = 0;
  while (sp->pore_fraction > 0 && (double)pore_vox < sp->pore_fraction * (double)n && tries < max_tries) {
    Sphere s = draw(sp->pore_rmin, sp->pore_rmax);
    tries++;
    if (!fits(s)) continue;
    place(s);
    pore_vox += stamp(s, sp->pore_label);
  }
  if (realised) {
    realised[0] = (double)incl_vox / (double)n;
    realised[1] = (double)pore_vox / (double)n;
  }
  return HPCR_OK;
}

}  // extern "C"
