/*
 * hpcrack_io.h — host-side input formats of HPCrack, kept byte-compatible:
 * config.txt (key=value), the gray-level material CSV and the uint8 .raw
 * image, plus the level geometry they imply and a seeded synthetic image
 * generator for benchmarks.  No device is touched by these functions.
 */
#ifndef HPCRACK_IO_H
#define HPCRACK_IO_H

#include <stddef.h>
#include "hpcrack.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Parameters (include/configuration.h:22-44).  `seen` has one bit per key in
 * the order of enum hpcr_config_key; unset keys keep the zero value. */
typedef struct {
  int D;
  int np[3];
  double X_start[3];
  char IMGname[1024];
  int voxel_nb[3];    /* voxel_nb_X/Y/Z */
  int roi_start[3];   /* ROI_start_X/Y/Z */
  double voxel_size;
  int element_nb[3];
  hpcr_mg mg_u, mg_d;
  double lc;          /* as parsed (an integer in the reference, configuration.cpp:72-73) */
  int outK, outG, outgc, outdesplacement, outH, outTdata, Tdatafrequence;
  unsigned long long seen;
} hpcr_config;

enum hpcr_config_key {
  HPCR_KEY_D = 0, HPCR_KEY_NP, HPCR_KEY_X_START, HPCR_KEY_IMGNAME, HPCR_KEY_VOXEL_NB_X, HPCR_KEY_VOXEL_NB_Y,
  HPCR_KEY_VOXEL_NB_Z, HPCR_KEY_ROI_X, HPCR_KEY_ROI_Y, HPCR_KEY_ROI_Z, HPCR_KEY_VOXEL_SIZE, HPCR_KEY_ELEMENT_NB,
  HPCR_KEY_NCY_U, HPCR_KEY_NU0_U, HPCR_KEY_NU1_U, HPCR_KEY_NU2_U, HPCR_KEY_GAMMA_U, HPCR_KEY_NCY_D,
  HPCR_KEY_NU0_D, HPCR_KEY_NU1_D, HPCR_KEY_NU2_D, HPCR_KEY_GAMMA_D, HPCR_KEY_LC, HPCR_KEY_OUTK, HPCR_KEY_OUTG,
  HPCR_KEY_OUTGC, HPCR_KEY_OUTDESPLACEMENT, HPCR_KEY_OUTH, HPCR_KEY_OUTTDATA, HPCR_KEY_TDATAFREQUENCE,
  HPCR_KEY_COUNT
};

/* Parameters::readFromFile (src/configuration.cpp:14-97).  strict != 0 keeps
 * the reference's parsing: the key is the raw text before '=', so lines
 * written "lc = 3.;" are ignored, and lc goes through stoi.  strict == 0
 * also trims blanks around the key and the value. */
int hpcr_read_config(const char* path, int strict, hpcr_config* out);

/* maxlevel (src/main.cpp:52-68), input_setting (src/image_input.cpp:21-29)
 * and initialize (src/initialization.cpp:19-47): the level stack a config
 * implies, and lc multiplied by the finest h (src/main.cpp:92-94). */
int hpcr_config_geometry(const hpcr_config* cfg, hpcr_geometry* geo, double* lc_scaled);

/* material_property (src/csv_reader.cpp:16-71): gray levels and E, nu, gc */
#define HPCR_MAX_PHASES 256
typedef struct {
  int n;
  int gray[HPCR_MAX_PHASES];
  double E[HPCR_MAX_PHASES], nu[HPCR_MAX_PHASES], gc[HPCR_MAX_PHASES];
} hpcr_material_table;
int hpcr_read_material_csv(const char* path, hpcr_material_table* out);

/* input / GL_MPI (src/image_input.cpp:33-118): the (nx,ny,nz) sub-block that
 * starts at roi_start of an x-fastest uint8 image of vx*vy*vz voxels. */
int hpcr_read_raw(const char* path, int vx, int vy, int vz, const int roi_start[3], int nx, int ny, int nz,
                  unsigned char* out);
int hpcr_write_raw(const char* path, const unsigned char* img, size_t n);

/* Load-step history files.
 * write_vector_3D / read_vector_3D (src/write_read_history.cpp:18-86,
 * include/write_read_history.h:17-20): one FP64 grid as raw doubles,
 * x-fastest.  A `name` containing ".dat" is the file path itself; otherwise
 * the file is <dir>/<name>_t_<t>_Proc_<proc>.dat (the reference's dir is
 * "Time_step_data").  Reading requires the file to hold exactly n doubles. */
int hpcr_write_vector_3d(const char* dir, const char* name, int t, int proc, const double* v, size_t n);
int hpcr_read_vector_3d(const char* dir, const char* name, int t, int proc, double* v, size_t n);

/* Scalar state of the load loop (crack_propagation, src/crack_propagation.cpp:
 * 28-123) at the end of step t, which the reference's history files do not
 * keep: the next load Ut, its increment, u_prop = Ut / U_unit, the frozen
 * f_norm of step 0 and Ue.  Text file <dir>/state_t_<t>.txt, one "key value"
 * per line, values printed with 17 significant digits. */
typedef struct hpcr_step_state {
  int t;
  double Ut, delta_U, u_prop, f_norm, Ue;
} hpcr_step_state;
int hpcr_write_step_state(const char* dir, const hpcr_step_state* s);
int hpcr_read_step_state(const char* dir, int t, hpcr_step_state* s);

/* image_MP (src/image_input.cpp:120-142): gray level -> K, G, gc with the
 * reference's integer-division interpolation; E is scaled by voxel_size.
 * Returns HPCR_ERR_ARG for a gray level outside the table. */
int hpcr_image_mp(const hpcr_material_table* tab, const unsigned char* img, size_t n, double voxel_size,
                  double* K, double* G, double* gc);

/* The material set-up of input_total (src/image_input.cpp:256-282) on the
 * device: the finest level's uint8 image (nodes of level maxlevel, x-fastest)
 * is uploaded (1 byte per node), image_MP runs on the device with results
 * identical to hpcr_image_mp, and coarse_VR fills K, G and gc on every coarse
 * level.  HPCR_ERR_ARG for a gray level outside the table. */
int hpcr_set_image(hpcr_ctx* ctx, const unsigned char* img, const hpcr_material_table* tab, double voxel_size);

/* Allocation-free variant of hpcr_set_image, with identical results.  It has
 * no counterpart in the reference: it is for callers that cannot allocate
 * device memory (serving loops that reuse pooled contexts, harnesses that
 * forbid hipMalloc).  It calls no hipMalloc/hipFree: the material table, the
 * out-of-table flag and a 4 MiB staging area through which the image streams
 * in chunks live in a fixed device scratch of about 4.01 MiB per device, held
 * until the process ends.  Limits: at most HPCR_MAX_PHASES (256) table rows,
 * else HPCR_ERR_ARG; any image size.  All calls on one device share that
 * scratch, so calls that overlap on one device (from different host threads)
 * race; calls one after another do not. */
int hpcr_set_image_noalloc(hpcr_ctx* ctx, const unsigned char* img, const hpcr_material_table* tab,
                           double voxel_size);

/* Seeded synthetic voxel image: non-overlapping spheres on integer node
 * centres placed by rejection (xoshiro256** seeded through splitmix64).
 * Inclusions first (count n_incl, radii U(incl_rmin, incl_rmax)), then pores
 * (radii U(pore_rmin, pore_rmax)) until their voxel fraction reaches
 * pore_fraction.  `gap` voxels separate any two spheres.  realised[0] and
 * realised[1] receive the inclusion and pore volume fractions. */
typedef struct {
  int dims[3];
  unsigned long long seed;
  int matrix_label;
  int n_incl;
  double incl_rmin, incl_rmax;
  int incl_label;
  double pore_fraction;
  double pore_rmin, pore_rmax;
  int pore_label;
  int gap;
  int max_tries;
} hpcr_synth_spec;
int hpcr_synth_image(const hpcr_synth_spec* spec, unsigned char* out, double realised[2]);

#ifdef __cplusplus
}
#endif
#endif /* HPCRACK_IO_H */
