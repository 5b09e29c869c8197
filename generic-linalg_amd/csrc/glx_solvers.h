// glx_solvers.h -- device-resident Krylov solvers (host orchestration API).
//
// Every solver keeps its vectors in HBM and its scalars in a small device
// state block that the reduction finalisers update; the host only enqueues
// kernels and, once per chunk of iterations, reads back the convergence word.
// Kernels check that word on entry and return at once after convergence, so
// the iteration sequence (and hence the iteration count) is exactly the
// reference's.
#pragma once

#include "glx_runtime.h"

#include <string>
#include <vector>

#include <complex>

namespace glx {

enum Verb : int { VERB_PASS_THROUGH = -1, VERB_NONE = 0, VERB_SUMMARY = 1, VERB_RESTART_DETAIL = 2, VERB_DETAIL = 3 };

// Result of one solve, mirroring inversion_info (inverter_struct.h:14-89).
struct SolveResult {
  double resSq = 0.0;
  int iter = 0;
  bool success = false;
  int ops_count = 0;
  std::string name;
  std::vector<double> relres_hist;  // iterated relative residual per iteration
  std::vector<double> resSqmrhs;    // multishift: per-shift true residual^2
  double bsqrt = 0.0;
  // mixed precision extras
  int reliable_updates = 0;
};

struct SolveOpts {
  int max_iter = 10000;
  double eps = 1e-10;
  bool want_hist = false;   // copy the per-iteration residuals back
  hipStream_t stream = nullptr;
  int chunk = 16;           // iterations enqueued between convergence checks
  int group = 8;            // single-pass CG: iterations per enqueued group
  bool use_graph = true;    // single-pass CG: replay one captured hipGraph per group
};

// x (in: initial guess, out: solution) and b are device pointers of the
// operator's element type (double for the real Laplacian, interleaved
// complex otherwise).
SolveResult dev_cg(const DevOperator& op, void* x, const void* b, const SolveOpts& o);
SolveResult dev_bicgstab(const DevOperator& op, void* x, const void* b, const SolveOpts& o);
SolveResult dev_gcr(const DevOperator& op, void* x, const void* b, const SolveOpts& o);
// Multishift CG: xs[n_shift] device solution vectors; shifts are read only
// (the reference's in-place permutation is kept inside the device state and
// undone before return).
SolveResult dev_cg_m(const DevOperator& op, void* const* xs, const void* b, int n_shift, const double* shifts,
                     int resid_freq_check, bool worst_first, const SolveOpts& o);

// Many independent multishift CG solves advanced together (batched CG-M; no
// reference counterpart as a batch, each system follows generic_cg_m.cpp):
// op = a staggered normal operator (its own lattice, links and mass), b its
// right-hand side, xs its n_shift solution vectors (device).  All systems use
// the same shifts.  Each phase of an iteration is one launch over all
// unconverged systems; groups of iterations replay one captured hipGraph
// when o.use_graph.
struct BatchSystem {
  DevOperator op;
  const void* b;
  void* const* xs;
};
// Work and time of a batched solve, per phase of the iteration: the D^dag D
// launches (96 B per site: r, the last p and 32 B of links read, p and Ap
// written; 64 B in the first iteration) and the residual pass (48 B: r read
// and written, Ap read).  bytes_shift counts the windows' x_s and p_s
// updates, which the residual pass's phase applies in a launch of its own
// (16 B per window term, 48 per x_s and 16 per p_s update); ms_shift stays
// 0.  The ms_* fields are the
// HIP-event time of each phase's launches when time_phases was asked for
// (direct launches, no graph), else 0.
struct BatchStats {
  double bytes_stencil = 0, bytes_r = 0, bytes_shift = 0;
  double ms_stencil = 0, ms_r = 0, ms_shift = 0;
  long long iterations = 0, upd_x = 0, upd_p = 0;
  int launches_shift = 0;
};
std::vector<SolveResult> dev_cg_m_batched(const std::vector<BatchSystem>& sys, int n_shift, const double* shifts,
                                          int resid_freq_check, bool worst_first, const SolveOpts& o,
                                          BatchStats* stats = nullptr, bool time_phases = false);

// Mixed-precision CG with reliable updates (no reference counterpart):
// FP32 Krylov vectors and FP32 or FP16 links, FP64 true residual and
// solution accumulation whenever the iterated residual drops by `delta`.
struct MixedOpts {
  int link_prec = LINKS_F32;
  double delta = 0.1;
  // 2: two CG iterations per launch (s-step form) where the operator is the
  // full-lattice D^dag D on a lattice that fits its strips; 1: one per
  // launch; 0: the process default (set_sstep, initially 2)
  int sstep = 0;
};

// Process-wide default of MixedOpts::sstep for the ABI's mixed solves, and the
// rows per block of the s-step kernel (a multiple of 4, >= 8) for every
// lattice size, or 0: per size (the table below, else sstep_rows_auto).
void set_sstep(int s, int rows);
int sstep_default();
int sstep_rows();
// Per lattice size: rows per block of the s-step kernel, or 0 for the
// one-iteration form at that size (rows < 0 removes the entry).
void set_sstep_rows(int lx, int ly, int rows);
int sstep_rows_for(int lx, int ly);
int sstep_rows_entry(int lx, int ly);  // the size's table entry, or -1 when it has none
int sstep_rows_auto(int lx, int ly);
// The form of the s-step launch on the full-lattice D^dag D: 0 blocks of four
// waves marching a strip together through LDS rings, 1 every wave marching a
// strip of its own in registers (band_rows rows per wave, 0: by size); -1
// (default): per lattice size, from the table (set_sstep_form_entry, the
// autotuner), else by size (sstep_form_for).
void set_sstep_form(int form, int band_rows);
int sstep_form();
void set_sstep_form_entry(int lx, int ly, int form);  // -1 removes the entry
int sstep_form_entry(int lx, int ly);                 // -1: none
int sstep_form_for(int lx, int ly);
int sw_rows_for(int lx, int ly);
double time_sscg_traffic_impl(const DevOperator& o, int link_prec, int reps);
SolveResult dev_cg_mixed(const DevOperator& op, void* x, const void* b, const SolveOpts& o, const MixedOpts& m);

// FP64 CG in the single-pass form the mixed solver uses (one fused stencil
// per iteration, |r_{k+1}|^2 from the recurrence): the same algorithm as
// dev_cg with a different rounding, for callers that want the fastest FP64
// solve rather than the reference's exact sequence of operations.
SolveResult dev_cg_fused(const DevOperator& op, void* x, const void* b, const SolveOpts& o);

// Live timing (HIP events) of the single-pass CG's fused iteration kernel on
// the staggered normal operator: prec 0 = FP64, LINKS_F32 / LINKS_F16 = the
// mixed solver's FP32 kernel with that link storage, 2 + LINKS_F32 /
// 2 + LINKS_F16 = the s-step kernel (two iterations per launch).  Returns ms
// per launch and the algorithmic HBM bytes per lattice site of one launch.
double time_cg_kernel(const DevOperator& op, int prec, int reps, double* bytes_per_site);

// Gradient of S(theta, m) = phi^dag (D^dag D)^-1 phi (no reference
// counterpart): one solve (precision 0: FP64 single-pass CG; LINKS_F32 /
// LINKS_F16: mixed-precision CG), psi = D chi, one per-link force pass.
// normal = the staggered normal operator, dop = D, both on the same links.
// grad: device, 2V doubles in the links' layout; chi_out: device or null.
struct GradResult {
  double S = 0.0;         // Re phi^dag chi
  double S_robust = 0.0;  // 2 Re phi^dag chi - |D chi|^2 (error second order in the residual)
  double dS_dm = 0.0;     // -2 m |chi|^2
  SolveResult solve;
};
GradResult dev_staggered_grad(const DevOperator& normal, const DevOperator& dop, const void* phi, int precision,
                              const SolveOpts& o, double* grad, void* chi_out);

// Even/odd solve of D x = b for the staggered operator on compacted even-site
// vectors (operators.cpp:543-616 keep full-length vectors): prepare
// b_e' = m b_e - D_eo b_o, CG on m^2 - D_eo D_oe (method 0: reference-order
// CG; 3: single-pass FP64 CG; LINKS_F32 / LINKS_F16: mixed-precision CG),
// reconstruct x_o = (b_o - D_oe x_e) / m.  x, b: full-lattice device
// vectors; the result describes the even-site solve; full_resSq (host, may
// be null) receives |D x - b|^2.
SolveResult dev_solve_staggered_eo(const DevOperator& op, void* x, const void* b, int method, const SolveOpts& o,
                                   double* full_resSq);
void dev_eo_prepare(const DevOperator& op, void* be, const void* b, hipStream_t st);
void dev_eo_reconstruct(const DevOperator& op, void* x, const void* xe, const void* b, hipStream_t st);

// Gauge fields on the device: Philox4x32-10 Gaussian (beta > 0) or uniform
// (beta = 0) link phases, and the mean plaquette.
void dev_gauss_links(void* u, int lx, int ly, double beta, unsigned long long seed, unsigned stream, hipStream_t st);
std::complex<double> dev_plaquette(const void* u, int lx, int ly, hipStream_t st);

// Plain operator application on device vectors.
void dev_apply(const DevOperator& op, void* out, const void* in, int link_prec, int fp32, hipStream_t st);

}  // namespace glx
