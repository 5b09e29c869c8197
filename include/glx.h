/*
 * glx.h -- C ABI of the MI355X-native (gfx950, HIP) solver path of
 * generic-linalg: Krylov solvers (CG, CG(n), multishift CG, BiCGStab,
 * BiCGStab(n), GCR, GCR(n)), the 2D Laplace and U(1) staggered operators
 * they call, and a mixed-precision CG with reliable updates.
 *
 * Plain C types only: complex<double> vectors are passed as double* pointing
 * at interleaved (re, im) pairs, the layout of std::complex<double>.  Every
 * function returns a glx_status (0 = success); glx_last_error() describes the
 * last failure of the calling thread.
 *
 * Two families of entry points:
 *   glx_minv_vector_*   host pointers, the reference's semantics one for one
 *                       (iteration count, success flag, ops_count, resSq,
 *                       verbosity lines).  When `matrix_vector` is one of the
 *                       registered device operators below, the whole solve
 *                       runs on the GPU; any other callback is called on the
 *                       host for each operator application, with the Krylov
 *                       vectors and reductions still on the device.
 *   glx_dev_*           device pointers (data already resident in HBM), for
 *                       callers that keep fields on the GPU across solves.
 *
 * Nothing here falls back to a CPU implementation: without a usable GPU the
 * functions return GLX_ERR_HIP.
 */
#ifndef GLX_H
#define GLX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GLX_ABI_VERSION 1

typedef enum {
  GLX_OK = 0,
  GLX_ERR_ARG = 1,      /* invalid argument */
  GLX_ERR_HIP = 2,      /* HIP runtime failure (no device, out of memory, ...) */
  GLX_ERR_INTERNAL = 3
} glx_status;

/* Levels of inversion_verbose_level (verbosity.h:9-16). */
enum {
  GLX_VERB_PASS_THROUGH = -1,
  GLX_VERB_NONE = 0,
  GLX_VERB_SUMMARY = 1,
  GLX_VERB_RESTART_DETAIL = 2,
  GLX_VERB_DETAIL = 3
};

/* inversion_verbose_struct (verbosity.h:18-24) as plain data.  The solver
 * prints the reference's lines (verbosity.h:73-129) to std::cout with its
 * current formatting, or, when out_buf is set, appends them to out_buf
 * (NUL-terminated, out_len updated) using float_format: 0 = "%g",
 * 1 = "%.<float_precision>e". */
typedef struct {
  int verbosity;
  const char* verb_prefix;
  int precond_verbosity;
  const char* precond_verb_prefix;
  char* out_buf;
  size_t out_cap;
  size_t out_len;
  int float_format;
  int float_precision;
} glx_verbose;

/* inversion_info (inverter_struct.h:14-89) as plain data.  resSqmrhs, when
 * non-null, is a caller buffer of n_rhs doubles that multishift solves fill. */
typedef struct {
  double resSq;
  int iter;
  int success;
  int ops_count;
  char name[32];
  double* resSqmrhs;
  int n_rhs;
  int reliable_updates; /* mixed-precision CG only */
} glx_inversion_info;

/* staggered_u1_op (operators.h:7-15), same field order.  lattice points at
 * 2*x_fine*y_fine complex links, AoS lattice[2*(y*x_fine + x) + mu]. */
typedef struct {
  double* lattice;
  double mass;
  int x_fine;
  int y_fine;
  int Nc;
  double wilson_coeff;
} glx_staggered_u1_op;

/* The operator callback contract of the reference (generic_cg.h:18-19):
 * lhs = A rhs, overwriting all `size` entries of lhs. */
typedef void (*glx_matvec)(double* lhs, double* rhs, void* extra_info);

/* ------------------------------------------------------------------------
 * Registered device operators.  Passing one of these as `matrix_vector`
 * runs the solve on the GPU; calling one directly applies the operator on
 * the GPU to host vectors.  extra_info is a glx_staggered_u1_op*.
 * ---------------------------------------------------------------------- */
/* real (double) Laplacian, diag 4 + mass, Nc colours
 * (tests/multishift/multishift.cpp:634-672) */
void glx_square_laplace_real(double* lhs, double* rhs, void* extra_info);
void glx_square_laplace(double* lhs, double* rhs, void* extra_info);              /* operators.cpp:28-66 */
void glx_square_laplace_u1(double* lhs, double* rhs, void* extra_info);           /* operators.cpp:73-120 */
void glx_square_staggered_u1(double* lhs, double* rhs, void* extra_info);         /* operators.cpp:184-239 */
void glx_gamma_5(double* lhs, double* rhs, void* extra_info);                     /* operators.cpp:242-259 */
void glx_square_staggered_gamma5_u1(double* lhs, double* rhs, void* extra_info);  /* operators.cpp:316-368 */
void glx_square_staggered_dagger_u1(double* lhs, double* rhs, void* extra_info);  /* operators.cpp:372-441 */
void glx_square_staggered_normal_u1(double* lhs, double* rhs, void* extra_info);  /* operators.cpp:444-453 */
void glx_square_staggered_deo_u1(double* lhs, double* rhs, void* extra_info);     /* operators.cpp:456-496 */
void glx_square_staggered_doe_u1(double* lhs, double* rhs, void* extra_info);     /* operators.cpp:499-539 */
void glx_square_staggered_m2mdeodoe_u1(double* lhs, double* rhs, void* extra_info); /* operators.cpp:566-588 */
/* the operators of operators.h off the solver path (operators.cpp:127-846),
 * applied on the GPU as general stencils; wilson_coeff scales the two-link
 * Laplacian term of the 2linklaplace operator */
void glx_square_staggered(double* lhs, double* rhs, void* extra_info);
void glx_square_staggered_gamma5(double* lhs, double* rhs, void* extra_info);
void glx_square_staggered_2linklaplace_u1(double* lhs, double* rhs, void* extra_info);
void glx_staggered_symmshift_x(double* lhs, double* rhs, void* extra_info);
void glx_staggered_symmshift_y(double* lhs, double* rhs, void* extra_info);
void glx_staggered_index_operator(double* lhs, double* rhs, void* extra_info);
/* operators.cpp:543-563 and :592-616 */
void glx_square_staggered_eoprec_prepare(double* rhs_e, double* rhs_orig, void* extra_info);
void glx_square_staggered_eoprec_reconstruct(double* lhs_full, double* lhs_e, double* rhs_o, void* extra_info);

/* stencil_2d (coarse_stencil.h:33-154) for Nc = 1 as plain data: clover[V],
 * hopping[4V] with blocks {+x, +y, -x, -y}, complex shifts as (re, im). */
typedef struct {
  int lx, ly;
  double* clover;
  double* hopping;
  double shift[2];
  double eo_shift[2];
  double dof_shift[2];
} glx_stencil_2d;
/* builders (operators_stencil.cpp:14-169); clover and hopping must point at
 * caller storage of V and 4V complex values */
int glx_get_square_staggered_u1_stencil(glx_stencil_2d* st, const glx_staggered_u1_op* op);
int glx_get_square_staggered_gamma5_u1_stencil(glx_stencil_2d* st, const glx_staggered_u1_op* op);
int glx_get_square_staggered_dagger_u1_stencil(glx_stencil_2d* st, const glx_staggered_u1_op* op);
/* operators with extra_info = glx_stencil_2d* (coarse_stencil.cpp:12-722,
 * operators_stencil.cpp:179-236) */
void glx_apply_stencil_2d(double* lhs, double* rhs, void* stencil);
void glx_apply_stencil_2d_eo(double* lhs, double* rhs, void* stencil);
void glx_apply_stencil_2d_oe(double* lhs, double* rhs, void* stencil);
void glx_apply_square_staggered_m2mdeodoe_stencil(double* lhs, double* rhs, void* stencil);
void glx_apply_square_staggered_eoprec_prepare_stencil(double* rhs_e, double* rhs_orig, void* stencil);
void glx_apply_square_staggered_eoprec_reconstruct_stencil(double* lhs_full, double* lhs_e, double* rhs_o, void* stencil);

/* stencil_2d for any number of colours nc, with optional two-link terms
 * (coarse_stencil.h:14-154), as plain data.  Coefficient (site s, row r,
 * column c) of a block is at (s*nc + r)*nc + c (the reference's layout):
 * clover[V nc^2]; hopping[4 V nc^2], blocks {+x, +y, -x, -y}; two_link
 * [8 V nc^2], blocks {+2x, +x+y, +2y, -x+y, -2x, -x-y, -2y, +x-y}, read only
 * when has_two.  sdir is the reference's stencil_dir (GLX_SDIR_*).  Vectors
 * have lx*ly*nc complex entries, entry (y*lx + x)*nc + colour. */
enum {
  GLX_SDIR_ALL = 0, GLX_SDIR_CLOVER = 1, GLX_SDIR_XP1 = 2, GLX_SDIR_YP1 = 3, GLX_SDIR_XM1 = 4, GLX_SDIR_YM1 = 5,
  GLX_SDIR_XP2 = 6, GLX_SDIR_XP1YP1 = 7, GLX_SDIR_YP2 = 8, GLX_SDIR_XM1YP1 = 9, GLX_SDIR_XM2 = 10,
  GLX_SDIR_XM1YM1 = 11, GLX_SDIR_YM2 = 12, GLX_SDIR_XP1YM1 = 13
};
typedef struct {
  int lx, ly, nc;
  int has_two;
  int sdir;
  double* clover;
  double* hopping;
  double* two_link;
  double shift[2];
  double eo_shift[2];
  double dof_shift[2];
} glx_stencil_2d_nc;
/* operators with extra_info = glx_stencil_2d_nc*; solvers given them run on
 * the GPU with the coefficients uploaded once per solve */
void glx_apply_stencil_2d_nc(double* lhs, double* rhs, void* stencil);     /* coarse_stencil.cpp:12-392 */
void glx_apply_stencil_2d_nc_eo(double* lhs, double* rhs, void* stencil);  /* coarse_stencil.cpp:395-557 */
void glx_apply_stencil_2d_nc_oe(double* lhs, double* rhs, void* stencil);  /* coarse_stencil.cpp:560-722 */
void glx_apply_stencil_2d_nc_tb(double* lhs, double* rhs, void* stencil);  /* coarse_stencil.cpp:725-1117 */
void glx_apply_stencil_2d_nc_bt(double* lhs, double* rhs, void* stencil);  /* coarse_stencil.cpp:1120-1512 */
/* operators_stencil.cpp:179-236 on any nc */
void glx_apply_stencil_2d_nc_m2mdeodoe(double* lhs, double* rhs, void* stencil);
void glx_stencil_2d_nc_eoprec_prepare(double* rhs_e, double* rhs_orig, void* stencil);
void glx_stencil_2d_nc_eoprec_reconstruct(double* lhs_full, double* lhs_e, double* rhs_o, void* stencil);
/* generate_stencil_2d (coarse_stencil.cpp:1515-1627): applies matrix_vector to
 * every unit vector and adds the entries into the (caller-zeroed) coefficient
 * arrays; a registered device operator runs on the GPU */
int glx_generate_stencil_2d_nc(glx_stencil_2d_nc* st, glx_matvec matrix_vector, void* extra_info);

/* Operator kinds (numbering of the device kernels). */
enum {
  GLX_OP_LAPLACE_REAL = 0,
  GLX_OP_LAPLACE = 1,
  GLX_OP_LAPLACE_U1 = 2,
  GLX_OP_STAGGERED = 3,
  GLX_OP_STAGGERED_DAGGER = 4,
  GLX_OP_STAGGERED_NORMAL = 5,
  GLX_OP_STAGGERED_GAMMA5 = 6,
  GLX_OP_GAMMA5 = 7,
  GLX_OP_DEO = 8,
  GLX_OP_DOE = 9,
  GLX_OP_M2MDEODOE = 10,
  GLX_OP_STENCIL = 11,
  GLX_OP_STENCIL_EO = 14,
  GLX_OP_STENCIL_OE = 15,
  GLX_OP_STENCIL_M2MDEODOE = 16,
  /* m^2 - D_eo D_oe (operators.cpp:566-588) on COMPACTED even-site vectors of
   * lx*ly/2 elements, site (x, y) with x + y even at y*(lx/2) + (x >> 1);
   * device API only (glx_dev_*); needs (lx/2) % 64 == 0 (FP64) or % 128 == 0
   * (FP32 vectors), ly % 4 == 0, ly >= 8 */
  GLX_OP_STAGGERED_EO_SCHUR = 17,
  /* general stencil_2d (glx_stencil_2d_nc): full, eo, oe, tb, bt pieces and
   * m^2 - D_eo D_oe */
  GLX_OP_CSTENCIL = 18,
  GLX_OP_CSTENCIL_EO = 19,
  GLX_OP_CSTENCIL_OE = 20,
  GLX_OP_CSTENCIL_TB = 21,
  GLX_OP_CSTENCIL_BT = 22,
  GLX_OP_CSTENCIL_M2MDEODOE = 23,
  /* the other operators of operators.h (extra_info = glx_staggered_u1_op*),
   * run as general stencils built on the device */
  GLX_OP_STAGGERED_FREE = 24,          /* square_staggered, operators.cpp:127-178 */
  GLX_OP_STAGGERED_GAMMA5_FREE = 25,   /* square_staggered_gamma5, :262-313 */
  GLX_OP_STAGGERED_2LINKLAPLACE = 26,  /* square_staggered_2linklaplace_u1, :626-694 */
  GLX_OP_SYMMSHIFT_X = 27,             /* staggered_symmshift_x, :697-736 */
  GLX_OP_SYMMSHIFT_Y = 28,             /* staggered_symmshift_y, :739-779 */
  GLX_OP_STAGGERED_INDEX = 29          /* staggered_index_operator, :782-846 */
};

/* Bind a caller's own host callback to a device operator, so that solves
 * given that callback run on the GPU (e.g. square_laplace.cpp:182-222, whose
 * extra_info is NULL: pass its lattice size and diagonal shift here).  With
 * params == NULL, extra_info of each solve is read as glx_staggered_u1_op*. */
int glx_register_operator(glx_matvec callback, int kind, const glx_staggered_u1_op* params);
int glx_unregister_operator(glx_matvec callback);

/* ------------------------------------------------------------------------
 * Solvers, host pointers (generic_cg.h:18-23, generic_cg_m.h:16-17,
 * generic_bicgstab.h:17-21, generic_gcr.h:17-22).  _r: real double vectors;
 * _c: complex vectors as interleaved doubles.  verbosity may be NULL.
 * ---------------------------------------------------------------------- */
int glx_minv_vector_cg_r(double* phi, double* phi0, int size, int max_iter, double eps, glx_matvec matrix_vector,
                         void* extra_info, glx_verbose* verbosity, glx_inversion_info* info);
int glx_minv_vector_cg_c(double* phi, double* phi0, int size, int max_iter, double eps, glx_matvec matrix_vector,
                         void* extra_info, glx_verbose* verbosity, glx_inversion_info* info);
int glx_minv_vector_cg_restart_r(double* phi, double* phi0, int size, int max_iter, double eps, int restart_freq,
                                 glx_matvec matrix_vector, void* extra_info, glx_verbose* verbosity,
                                 glx_inversion_info* info);
int glx_minv_vector_cg_restart_c(double* phi, double* phi0, int size, int max_iter, double eps, int restart_freq,
                                 glx_matvec matrix_vector, void* extra_info, glx_verbose* verbosity,
                                 glx_inversion_info* info);
int glx_minv_vector_bicgstab_r(double* phi, double* phi0, int size, int max_iter, double eps,
                               glx_matvec matrix_vector, void* extra_info, glx_verbose* verbosity,
                               glx_inversion_info* info);
int glx_minv_vector_bicgstab_c(double* phi, double* phi0, int size, int max_iter, double eps,
                               glx_matvec matrix_vector, void* extra_info, glx_verbose* verbosity,
                               glx_inversion_info* info);
int glx_minv_vector_bicgstab_restart_r(double* phi, double* phi0, int size, int max_iter, double eps,
                                       int restart_freq, glx_matvec matrix_vector, void* extra_info,
                                       glx_verbose* verbosity, glx_inversion_info* info);
int glx_minv_vector_bicgstab_restart_c(double* phi, double* phi0, int size, int max_iter, double eps,
                                       int restart_freq, glx_matvec matrix_vector, void* extra_info,
                                       glx_verbose* verbosity, glx_inversion_info* info);
int glx_minv_vector_gcr_r(double* phi, double* phi0, int size, int max_iter, double eps, glx_matvec matrix_vector,
                          void* extra_info, glx_verbose* verbosity, glx_inversion_info* info);
int glx_minv_vector_gcr_c(double* phi, double* phi0, int size, int max_iter, double eps, glx_matvec matrix_vector,
                          void* extra_info, glx_verbose* verbosity, glx_inversion_info* info);
int glx_minv_vector_gcr_restart_r(double* phi, double* phi0, int size, int max_iter, double eps, int restart_freq,
                                  glx_matvec matrix_vector, void* extra_info, glx_verbose* verbosity,
                                  glx_inversion_info* info);
int glx_minv_vector_gcr_restart_c(double* phi, double* phi0, int size, int max_iter, double eps, int restart_freq,
                                  glx_matvec matrix_vector, void* extra_info, glx_verbose* verbosity,
                                  glx_inversion_info* info);
/* The reference's dispatcher minv_unpreconditioned (generic_inverter.cpp:18-188,
 * generic_inverters.h:70-111): type is the reference's minv_inverter value
 * (MINV_CG 0, MINV_GCR 2, MINV_BICGSTAB 3 run here, restarted when
 * params->restart; MINV_INVALID -1 returns a zeroed info; the reference's
 * other solvers -- CR 1, BiCGStab(l) 4, GMRES 5, SOR 6, MinRes 7 -- are not
 * part of this library and return GLX_ERR_ARG).  glx_inverter_params mirrors
 * minv_inverter_params field by field. */
typedef struct {
  double tol;
  int max_iters;
  int restart;
  int restart_freq;
  double sor_omega;
  double minres_omega;
  int bicgstabl_l;
} glx_inverter_params;
int glx_minv_unpreconditioned_r(double* lhs, double* rhs, int size, int type, const glx_inverter_params* params,
                                glx_matvec matrix_vector, void* extra_info, glx_verbose* verbosity,
                                glx_inversion_info* info);
int glx_minv_unpreconditioned_c(double* lhs, double* rhs, int size, int type, const glx_inverter_params* params,
                                glx_matvec matrix_vector, void* extra_info, glx_verbose* verbosity,
                                glx_inversion_info* info);
/* Multishift CG: phi[n_shift] host vectors (zeroed, then solutions);
 * shifts[n_shift] is read (left unchanged on return, as the reference
 * restores it); info->resSqmrhs (n_shift doubles) receives the per-shift
 * squared true residuals; info->resSq is 0 as in the reference. */
int glx_minv_vector_cg_m_r(double** phi, double* phi0, int n_shift, int size, int resid_freq_check, int max_iter,
                           double eps, double* shifts, glx_matvec matrix_vector, void* extra_info, int worst_first,
                           glx_verbose* verbosity, glx_inversion_info* info);
int glx_minv_vector_cg_m_c(double** phi, double* phi0, int n_shift, int size, int resid_freq_check, int max_iter,
                           double eps, double* shifts, glx_matvec matrix_vector, void* extra_info, int worst_first,
                           glx_verbose* verbosity, glx_inversion_info* info);

/* ------------------------------------------------------------------------
 * Device-resident API.
 * ---------------------------------------------------------------------- */
/* A gauge field resident in HBM: FP64 links, plus FP32 and FP16 (+ per-site
 * FP32 scale) copies when flags include GLX_LINKS_F32 / GLX_LINKS_F16.
 * links is a host array of 2*lx*ly complex values (AoS as above). */
enum { GLX_LINKS_F32 = 1, GLX_LINKS_F16 = 2 };
typedef struct glx_lattice_s* glx_lattice;
int glx_lattice_create(glx_lattice* lat, const double* links, int lx, int ly, int flags);
int glx_lattice_update_links(glx_lattice lat, const double* links);
int glx_lattice_destroy(glx_lattice lat);
/* Gauge fields generated and measured in HBM: the device counterparts of
 * gauss_gauge_u1 / rand_gauge_u1 and get_plaquette_u1 (u1_utils.cpp:79-112,
 * 190-207), for setting up many lattices without host traffic.  Each site s
 * draws from the counter-based Philox4x32-10 generator (counter
 * {s, stream, 0, 0}, key = seed), so a field depends only on (seed, stream).
 * theta ~ N(0, 1/|beta|) by Box-Muller; beta = 0 gives uniform phases on
 * [-pi, pi) (a hot start).  The host std::mt19937 stream of the reference is
 * not reproduced: for fields equal to the reference's, generate them on the
 * host (compat/u1_utils.h) and use glx_lattice_create. */
int glx_lattice_create_gaussian(glx_lattice* lat, int lx, int ly, int flags, double beta, unsigned long long seed,
                                unsigned int stream);
int glx_lattice_gaussian_links(glx_lattice lat, double beta, unsigned long long seed, unsigned int stream);
/* Host copy of the FP64 links (2*lx*ly complex values). */
int glx_lattice_get_links(glx_lattice lat, double* links);
/* Mean plaquette, plaq[0] + i plaq[1] (get_plaquette_u1, u1_utils.cpp:190-207). */
int glx_lattice_plaquette(glx_lattice lat, double* plaq);

/* Solver selection for glx_dev_solve. */
enum { GLX_SOLVER_CG = 0, GLX_SOLVER_BICGSTAB = 1, GLX_SOLVER_GCR = 2 };
/* x_dev (in: guess, out: solution) and b_dev are device vectors of
 * lx*ly*nc elements (double for GLX_OP_LAPLACE_REAL, complex otherwise).
 * lat may be NULL for operators without links. */
int glx_dev_solve(int solver, glx_lattice lat, int kind, double mass, int nc, double* x_dev, const double* b_dev,
                  int max_iter, double eps, glx_verbose* verbosity, glx_inversion_info* info);
int glx_dev_cg_m(glx_lattice lat, int kind, double mass, int nc, double** x_dev, const double* b_dev, int n_shift,
                 const double* shifts, int resid_freq_check, int max_iter, double eps, int worst_first,
                 glx_verbose* verbosity, glx_inversion_info* info);
/* Batched multishift CG (no reference counterpart as a batch; each system
 * follows minv_vector_cg_m, generic_cg_m.cpp:312-600, on D^dag D of its own
 * lattice and mass): n_sys independent solves advanced together, every phase
 * of an iteration one launch over all unconverged systems, groups of
 * iterations replayed from one captured hipGraph when use_graph.  All systems
 * share n_shift and shifts.  x_dev of system s: n_shift device vectors
 * (zeroed, then the solutions, in the order of shifts).  infos: n_sys
 * records (may be NULL); a non-NULL resSqmrhs receives that system's
 * per-shift squared true residuals.  Lattices need lx % 64 == 0,
 * ly % 4 == 0, ly >= 8. */
typedef struct {
  glx_lattice lat;
  double mass;
  const double* b_dev;
  double** x_dev;
} glx_batch_system;
int glx_dev_cg_m_batched(int n_sys, const glx_batch_system* sys, int n_shift, const double* shifts,
                         int resid_freq_check, int max_iter, double eps, int worst_first, int use_graph,
                         glx_inversion_info* infos);
/* The same solve, also reporting its work per phase of an iteration as
 * algorithmic HBM bytes.  An iteration has two phases: the D^dag D launch,
 * which forms p = r + alpha p in its prologue (96 B per site: r, the last p
 * and 32 B of links read, p and Ap written; 64 B in the first iteration),
 * and the residual pass r += beta Ap (48 B per site).  The shifts' x_s and
 * p_s are updated window by window (up to 10 iterations at a time) in the
 * residual pass's phase, by a launch of its own before r += beta Ap:
 * bytes_shift_pass counts that work (16 B per site per window term read, 48
 * per x_s update, 16 per p_s update), and its time is part of ms_r_pass
 * (ms_shift_pass stays 0).  With time_phases the launches are issued
 * directly (no graph) between HIP events and ms_* is each phase's total
 * time; else ms_* = 0.  No reference counterpart. */
typedef struct {
  double bytes_stencil, bytes_r_pass, bytes_shift_pass;
  double ms_stencil, ms_r_pass, ms_shift_pass;
  long long iterations;     /* summed over the systems */
  long long shift_updates_x; /* x_s updates summed over systems and iterations */
  long long shift_updates_p; /* p_s updates */
} glx_batch_stats;
int glx_dev_cg_m_batched_ex(int n_sys, const glx_batch_system* sys, int n_shift, const double* shifts,
                            int resid_freq_check, int max_iter, double eps, int worst_first, int use_graph,
                            int time_phases, glx_inversion_info* infos, glx_batch_stats* stats);
/* Mixed-precision CG (no reference counterpart): FP32 Krylov vectors with
 * FP32 (link_prec = 1) or FP16 + scale (link_prec = 2) links; FP64 true
 * residual r = b - A x and FP64 solution accumulation at every reliable
 * update, triggered when the iterated residual falls below delta times its
 * maximum since the last update.  Stops when the FP64 true relative
 * residual is below eps. */
int glx_dev_cg_mixed(glx_lattice lat, int kind, double mass, double* x_dev, const double* b_dev, int max_iter,
                     double eps, double delta, int link_prec, glx_verbose* verbosity, glx_inversion_info* info);
/* glx_dev_cg_mixed without any allocation (no reference counterpart: for
 * callers that cannot allocate -- serving loops, harnesses that forbid
 * hipMalloc): the same solve, bit for bit, with every device buffer taken
 * from a fixed 112 MiB array per GPU that the library holds for the life of
 * the process, on the calling thread's per-thread default stream, creating
 * no stream, event, graph or pinned host memory.  At most 2^20 lattice sites
 * (1024 x 1024) and 2^18 iterations; larger sizes return GLX_ERR_ARG.  Calls
 * on one GPU share its array: calls that overlap on one GPU race. */
int glx_dev_cg_mixed_noalloc(glx_lattice lat, int kind, double mass, double* x_dev, const double* b_dev,
                             int max_iter, double eps, double delta, int link_prec, glx_verbose* verbosity,
                             glx_inversion_info* info);
/* FP64 CG in single-pass form (no reference counterpart as such): the same
 * algorithm as GLX_SOLVER_CG with one fused stencil launch per iteration
 * (|r_{k+1}|^2 from the recurrence), so the rounding -- not the method --
 * differs from the reference's sequence of operations. */
int glx_dev_cg_fused(glx_lattice lat, int kind, double mass, double* x_dev, const double* b_dev, int max_iter,
                     double eps, glx_verbose* verbosity, glx_inversion_info* info);
/* Gradient of S(theta, m) = phi^dag (D^dag D)^-1 phi for the U(1) staggered
 * operator of the lattice (no reference counterpart; the operator is
 * operators.cpp:444-453, links U = exp(i theta)): chi = (D^dag D)^-1 phi by
 * CG to eps (precision 0: FP64 single-pass CG; 1 / 2: mixed-precision CG with
 * FP32 / FP16 links), psi = D chi, then per link
 *   dS/dtheta_mu(x) = -eta_mu(x) Im[psi(x)^* U_mu(x) chi(x+mu) + psi(x+mu)^* U_mu(x)^* chi(x)]
 * into grad_dev[2*(y*lx + x) + mu] (2*lx*ly doubles, the links' layout) and
 * dS/dm = -2 m |chi|^2.  S = Re phi^dag chi; S_robust = 2 Re phi^dag chi -
 * |D chi|^2 differs from the exact S only at second order in the solve
 * error.  phi_dev: lx*ly complex; chi_dev (may be NULL) receives chi.
 * dS_dm, S, S_robust, info may be NULL. */
int glx_dev_staggered_grad(glx_lattice lat, double mass, const double* phi_dev, int precision, int max_iter,
                           double eps, double* grad_dev, double* chi_dev, double* dS_dm, double* S, double* S_robust,
                           glx_inversion_info* info);
/* Even/odd solve of D x = b for the staggered operator (the pipeline of
 * operators.cpp:543-616 -- prepare, CG on m^2 - D_eo D_oe, reconstruct -- on
 * compacted even-site vectors): x_dev, b_dev full-lattice complex vectors;
 * method 0: reference-order CG, 3: single-pass FP64 CG, 1 / 2: mixed
 * precision with FP32 / FP16 links.  info describes the even-site solve
 * (its eps is relative to |b_e'|); full_resSq (may be NULL) receives
 * |D x - b|^2.  Needs m != 0 and the GLX_OP_STAGGERED_EO_SCHUR sizes. */
int glx_dev_solve_staggered_eo(glx_lattice lat, double mass, double* x_dev, const double* b_dev, int method,
                               int max_iter, double eps, double* full_resSq, glx_inversion_info* info);
/* The two ends of that pipeline: be_dev (compact, lx*ly/2) = m b_e - D_eo b_o;
 * x_dev (full) = (x_e scattered, (b_o - D_oe x_e) / m on odd sites). */
int glx_dev_eo_prepare(glx_lattice lat, double mass, double* be_dev, const double* b_dev);
int glx_dev_eo_reconstruct(glx_lattice lat, double mass, double* x_dev, const double* xe_dev, const double* b_dev);
/* Launch tuning of the single-pass CG and the mixed-precision CG: iterations
 * per enqueued group (default 16; rounded up to even where the iteration is
 * self-finalising) and whether a group is captured once as a hipGraph and
 * replayed (default 0: measured slightly slower than direct launches). */
int glx_set_tuning(int group, int use_graph);
/* Iterations per launch of the mixed-precision CG on D^dag D (no reference
 * counterpart): s = 2 (default) runs two CG iterations per launch in s-step
 * form (basis p, Ap, A^2 p, r, Ar and its Gram matrix, FP32 vectors, FP64
 * coordinates) where lx % 128 == 0, ly % 4 == 0 and ly >= 8; s = 1 one
 * iteration per launch.  rows: rows per block of the s-step kernel for every
 * lattice size (a multiple of 4, >= 8), or 0 (default): per size, from the
 * table below, else the fewest of 8, 16, 32, 64, 128 that keep the launch
 * within 512 blocks of 128 lattice columns (256 beyond). */
int glx_set_sstep(int s, int rows);
int glx_get_sstep(int* s, int* rows);
/* The per-size table: rows per block of the s-step kernel for an lx x ly
 * lattice, or 0 = the one-iteration form at that size; -1 removes the entry.
 * glx_get_sstep_rows returns what a solve of that size would use. */
int glx_set_sstep_rows(int lx, int ly, int rows);
int glx_get_sstep_rows(int lx, int ly, int* rows);
/* The form of the s-step launch on the full-lattice D^dag D (no reference
 * counterpart; the results agree to the rounding of the hops and the Gram
 * sums): form 0 blocks of four waves march a strip of 128 columns together
 * through LDS rings; form 1 every wave marches a strip of 56 columns of its
 * own in registers, band_rows rows per wave (0: the fewest of 10, 16, ...,
 * 52 that keep the launch within 1024 waves; at most 56).  form -1 (default)
 * chooses per lattice size: the size's entry (glx_set_sstep_form_entry, set
 * by glx_autotune_sstep), else form 1 from 512^2 sites up to lattices that
 * fit 1024 waves at bands of 52 rows, form 0 otherwise.
 * glx_get_sstep_form_for returns what a solve of that size would use.  The
 * mixed CG of the even/odd solve (the compact Schur operator) follows the
 * same choice, except that by size it takes form 1 only above 512^2 sites. */
int glx_set_sstep_form(int form, int band_rows);
int glx_get_sstep_form(int* form);
int glx_set_sstep_form_entry(int lx, int ly, int form);
int glx_get_sstep_form_for(int lx, int ly, int* form);
/* Autotuner for the mixed-precision CG on D^dag D of the lattice: times
 * `reps` iterations in the one-iteration form, in s-step form 0 at 8, 16,
 * 32, 64 and 128 rows per block (and 256 where ly >= 2048) and in s-step
 * form 1, installs the fastest in the tables above (rows per block, and the
 * form for the size) and returns the rows (0 = one-iteration form) with the
 * time per iteration in ms; glx_get_sstep_form_for gives the form.  Needs
 * FP32 links.  No reference counterpart. */
int glx_autotune_sstep(glx_lattice lat, int reps, int* rows, double* ms_per_iteration);
/* Launch shape of the fused normal operator D^dag D (no reference
 * counterpart; operators.cpp:444-453 is the operator it computes): rows per
 * block of the row-marching kernel for an lx x ly lattice and vector element
 * size esize (16 = complex FP64, 8 = complex FP32); 0 selects the tiled
 * kernel.  lx = ly = 0 sets the default for sizes not in the table; the
 * initial default, -1, takes the fewest of 16/32/64/128 rows that keep the
 * grid within 512 blocks (64 lanes of 1 FP64 or 2 FP32 sites per block row;
 * compact rows of lx / 2 sites for the even/odd Schur operator), else 256.
 * rows = -1 with a size removes that size's entry.
 * glx_get_normal_rows returns what a solve of that size would use (for
 * lx = ly = 0, the default itself).
 * The row-marching kernel needs lx % 64 == 0 (FP64) or lx % 128 == 0 (FP32,
 * two sites per lane), ly % 4 == 0, ly >= 8; other sizes always use the
 * tiled kernel. */
int glx_set_normal_rows(int lx, int ly, int esize, int rows);
int glx_get_normal_rows(int lx, int ly, int esize, int* rows);
/* Launch shape of the plain staggered D, D^dag, gamma5 D, D_eo and D_oe on
 * complex vectors (no reference counterpart; operators.cpp:184-239,
 * :316-368, :372-441 and :456-539 are the operators they compute): rows per
 * wave of the wave-march kernel, which
 * serves lattices with lx % 64 == 0 and ly >= 4 (other sizes, and the
 * solvers' fused launches, use the tiled kernel).  rows = -1 (the initial
 * setting): by lattice size (16 rows for complex FP32, 18 for FP64 or 16 on
 * lattices that give 2048 waves, fewer where the lattice would fill fewer
 * than 768 SIMDs); 0: the tiled kernel
 * always; > 0: that many rows.  glx_get_dwave_rows returns what an lx x ly
 * apply with vector element size esize (8 = complex FP32, 16 = complex FP64)
 * would use (0: the tiled kernel). */
int glx_set_dwave_rows(int rows);
/* The same for the wave-march kernel of the plain fused D^dag D (no
 * reference counterpart; operators.cpp:444-453 is the operator), which
 * serves lattices with lx even, lx >= 64 and ly >= 4 (60 owned columns per
 * wave); the solvers' fused iterations keep the row-marching and tiled
 * kernels.  rows = -1 (the initial setting): by lattice size (none where
 * one application moves more than 256 MiB, else 24 rows where that gives
 * 2048 waves, 10 where that gives 768, else about one wave per SIMD); 0: the row-marching or tiled kernel always; > 0: that many. */
int glx_set_ddwave_rows(int rows);
int glx_get_ddwave_rows(int lx, int ly, int esize, int* rows);
int glx_get_dwave_rows(int lx, int ly, int esize, int* rows);
/* Autotuner: for every candidate launch shape (tiled; 8..128 rows per block,
 * and 256 where ly >= 2048)
 * times `reps` iterations of the single-pass FP64 CG (glx_dev_cg_fused) and,
 * if the lattice has FP32 links, of the mixed-precision CG in its
 * one-iteration form (glx_set_sstep(1, ...)) on D^dag D of the lattice, records the fastest shape per precision in the table above and
 * returns it with its time per iteration in ms.  Any output pointer may be
 * NULL. */
int glx_autotune_normal(glx_lattice lat, int reps, int* rows64, int* rows32, double* ms64, double* ms32);
/* out = A in on device vectors; fp32 selects complex<float> vectors with the
 * FP32 (link_prec 1) or FP16 (link_prec 2) links.  out and in must be
 * different buffers (every kernel reads neighbours of sites another wave
 * writes); the same pointer is an error. */
int glx_dev_apply(glx_lattice lat, int kind, double mass, int nc, int fp32, int link_prec, void* out_dev,
                  const void* in_dev);
/* Average duration in ms of `reps` back-to-back applications, timed with HIP
 * events on the stream the kernels run on. */
int glx_dev_time_apply(glx_lattice lat, int kind, double mass, int fp32, int link_prec, void* out_dev,
                       const void* in_dev, int reps, double* ms_per_apply);

/* The streaming ceiling the stencils are read against (no reference
 * counterpart): average ms per pass of a plain kernel over a footprint of
 * `bytes` (16-byte accesses per lane, HIP events around `reps` passes on the
 * null stream): mode 0 reads it, 1 writes it, 2 copies half of it into the
 * other half, 3 reads three quarters and writes the fourth (the stencils'
 * 3 : 1 ratio of bytes read to bytes written).  bytes_per_pass (may be NULL)
 * receives the bytes one pass moves. */
int glx_dev_time_stream(int mode, size_t bytes, int reps, double* ms_per_pass, double* bytes_per_pass);

/* Live timing (HIP events around `reps` back-to-back launches on one stream)
 * of the single-pass CG's fused iteration kernel on D^dag D of the lattice:
 * prec 0 = FP64 (glx_dev_cg_fused), 1 / 2 = the mixed-precision CG's FP32
 * kernel with FP32 / FP16 links, 3 / 4 = its s-step kernel (two iterations
 * per launch) with FP32 / FP16 links.  Returns ms per launch and the algorithmic
 * HBM bytes per site one launch moves.  No reference counterpart. */
int glx_dev_time_cg_kernel(glx_lattice lat, double mass, int prec, int reps, double* ms_per_launch,
                           double* bytes_per_site);
/* Live timing of the memory traffic of the wave-strip s-step launch alone:
 * the same waves, loads and stores as the launch glx_dev_cg_mixed runs on
 * this lattice in that form, without its arithmetic (link_prec
 * GLX_LINKS_F32 or GLX_LINKS_F16) -- the floor the access pattern sets for
 * that launch.  Returns ms per launch.  No reference counterpart. */
int glx_dev_time_sscg_traffic(glx_lattice lat, int link_prec, int reps, double* ms_per_launch);

/* A general stencil resident in HBM (coefficients transposed for coalesced
 * reads), for device vectors of lx*ly*nc complex entries.  kind is one of
 * GLX_OP_CSTENCIL*; sdir overrides the stencil's own (GLX_SDIR_*).  A
 * two-link sdir on a stencil without two-link terms leaves out_dev as it was
 * (full, tb and bt pieces), as the reference leaves lhs. */
typedef struct glx_cstencil_s* glx_cstencil;
int glx_cstencil_create(glx_cstencil* cs, const glx_stencil_2d_nc* st);
int glx_cstencil_destroy(glx_cstencil cs);
int glx_dev_cstencil_apply(glx_cstencil cs, int kind, int sdir, double* out_dev, const double* in_dev);
/* CG / BiCGStab / GCR (GLX_SOLVER_*) on device vectors with the stencil */
int glx_dev_cstencil_solve(int solver, glx_cstencil cs, int kind, double* x_dev, const double* b_dev, int max_iter,
                           double eps, glx_verbose* verbosity, glx_inversion_info* info);
/* ms per application (HIP events around reps back-to-back launches) */
int glx_dev_time_cstencil(glx_cstencil cs, int kind, double* out_dev, const double* in_dev, int reps,
                          double* ms_per_apply);

/* Device memory helpers (so that callers need no HIP headers). */
int glx_dev_alloc(void** ptr, size_t bytes);
int glx_dev_free(void* ptr);
int glx_memcpy_htod(void* dst_dev, const void* src_host, size_t bytes);
int glx_memcpy_dtoh(void* dst_host, const void* src_dev, size_t bytes);
int glx_memcpy_dtod(void* dst_dev, const void* src_dev, size_t bytes);
int glx_device_synchronize(void);
int glx_device_count(int* count);
/* Frees the library's cached device buffers. */
int glx_release_cache(void);
const char* glx_last_error(void);
int glx_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* GLX_H */
