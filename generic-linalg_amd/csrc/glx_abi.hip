// glx_abi.hip -- the C ABI (include/glx.h) over the device solvers.
//
// Host-pointer entry points keep the reference's contract
// (generic_cg.cpp:132-427, generic_cg_m.cpp:23-600, generic_bicgstab.cpp:22-385,
// generic_gcr.cpp:27-378): the operator callback is recognised by pointer
// identity and its extra_info (a staggered_u1_op or stencil_2d mirror) is
// turned into a device operator; vectors and links are copied to HBM, the
// solve runs on the GPU, and the solution is copied back.  Restarted solvers
// repeat the device solve on the resident vectors exactly as the reference's
// restart wrappers do, and every verbosity line is reproduced
// (verbosity.h:73-129).
#include "../../include/glx.h"
#include "glx_internal.h"
#include "glx_solvers.h"

#include <algorithm>
#include <chrono>
#include <complex>
#include <cmath>
#include <cstring>
#include <iostream>
#include <memory>
#include <mutex>
#include <sstream>
#include <unordered_map>
#include <vector>

using namespace glx;

// ================================================================ errors ==
static thread_local std::string g_last_error;
// launch tuning of the single-pass and mixed-precision CG (glx_set_tuning)
static int g_group = 16;
static int g_use_graph = 0;

template <class F>
static int guarded(F&& f) {
  try {
    f();
    return GLX_OK;
  } catch (const ArgError& e) {
    g_last_error = e.what();
    return GLX_ERR_ARG;
  } catch (const HipError& e) {
    g_last_error = e.what();
    return GLX_ERR_HIP;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return GLX_ERR_INTERNAL;
  }
}

// ===================================================== verbosity output ==
namespace {

struct Printer {
  glx_verbose* v;
  std::ostringstream line;
  explicit Printer(glx_verbose* vb) : v(vb) {
    if (v && !v->out_buf) {
      line.flags(std::cout.flags());
      line.precision(std::cout.precision());
    } else if (v && v->float_format == 1) {
      line << std::scientific;
      line.precision(v->float_precision > 0 ? v->float_precision : 6);
    }
  }
  int level() const { return v ? v->verbosity : GLX_VERB_NONE; }
  std::string prefix() const { return v && v->verb_prefix ? v->verb_prefix : ""; }
  void flush() {
    const std::string s = line.str();
    line.str(std::string());
    if (!v) return;
    if (v->out_buf) {
      if (v->out_cap == 0) return;
      size_t room = v->out_cap - 1 > v->out_len ? v->out_cap - 1 - v->out_len : 0;
      size_t n = std::min(room, s.size());
      std::memcpy(v->out_buf + v->out_len, s.data(), n);
      v->out_len += n;
      v->out_buf[v->out_len] = 0;
    } else {
      std::cout << s;
    }
  }
  // verbosity.h:73-84
  void resid(const std::string& alg, int iter, int ops, double relres) {
    if (level() != GLX_VERB_DETAIL) return;
    line << prefix() << alg << " Iter " << iter << " Ops " << ops << " RelRes " << relres << "\n";
    flush();
  }
  // verbosity.h:87-98
  void summary(const std::string& alg, bool success, int iter, int ops, double relres) {
    const int l = level();
    if (l != GLX_VERB_SUMMARY && l != GLX_VERB_RESTART_DETAIL && l != GLX_VERB_DETAIL) return;
    line << prefix() << alg << " Success " << (success ? "Y" : "N") << " Iter " << iter << " Ops " << ops
         << " RelRes " << relres << "\n";
    flush();
  }
  // verbosity.h:101-117 (the algorithm name is always "CG-M ")
  void summary_multi(bool success, int iter, int ops, const std::vector<double>& relres) {
    const int l = level();
    if (l != GLX_VERB_SUMMARY && l != GLX_VERB_RESTART_DETAIL && l != GLX_VERB_DETAIL) return;
    line << prefix() << "CG-M " << " Success " << (success ? "Y" : "N") << " Iter " << iter << " Ops " << ops
         << " RelRes ";
    for (double r : relres) line << r << " ";
    line << "\n";
    flush();
  }
  // verbosity.h:120-129
  void restart(const std::string& alg, int iter, int ops, double relres) {
    const int l = level();
    if (l != GLX_VERB_RESTART_DETAIL && l != GLX_VERB_DETAIL) return;
    line << prefix() << alg << " Iter " << iter << " Ops " << ops << " RelRes " << relres << "\n";
    flush();
  }
};

// verbosity.h:27-43: inner solves of a restarted solver print only at DETAIL
glx_verbose restart_verbosity(const glx_verbose* v) {
  glx_verbose r;
  std::memset(&r, 0, sizeof(r));
  if (v) {
    r = *v;
    if (v->verbosity == GLX_VERB_RESTART_DETAIL || v->verbosity == GLX_VERB_SUMMARY) r.verbosity = GLX_VERB_NONE;
  }
  return r;
}

// ================================================== operator registry ====
struct RegEntry {
  int kind;
  bool has_params;
  glx_staggered_u1_op params;
  int (*translate)(void*, glx_stencil_2d_nc*) = nullptr;  // general stencils given in another form
};
struct Registry {
  std::mutex mu;
  std::unordered_map<void*, RegEntry> map;
  static Registry& get() {
    static Registry* r = new Registry;
    return *r;
  }
  void add(void* f, int kind, const glx_staggered_u1_op* p, int (*tr)(void*, glx_stencil_2d_nc*) = nullptr) {
    std::lock_guard<std::mutex> g(mu);
    RegEntry e{kind, p != nullptr, {}, tr};
    if (p) e.params = *p;
    map[f] = e;
  }
  bool find(void* f, RegEntry& out) {
    std::lock_guard<std::mutex> g(mu);
    auto it = map.find(f);
    if (it == map.end()) return false;
    out = it->second;
    return true;
  }
  void remove(void* f) {
    std::lock_guard<std::mutex> g(mu);
    map.erase(f);
  }
};

bool is_stencil_kind(int k) { return op_coeffs(k); }
bool real_kind(int k) { return k == OP_LAPLACE_REAL; }

struct BuiltinInit {
  BuiltinInit() {
    auto& r = Registry::get();
    r.add(reinterpret_cast<void*>(&glx_square_laplace_real), OP_LAPLACE_REAL, nullptr);
    r.add(reinterpret_cast<void*>(&glx_square_laplace), OP_LAPLACE, nullptr);
    r.add(reinterpret_cast<void*>(&glx_square_laplace_u1), OP_LAPLACE_U1, nullptr);
    r.add(reinterpret_cast<void*>(&glx_square_staggered_u1), OP_STAG, nullptr);
    r.add(reinterpret_cast<void*>(&glx_square_staggered_dagger_u1), OP_STAG_DAG, nullptr);
    r.add(reinterpret_cast<void*>(&glx_square_staggered_normal_u1), OP_STAG_NORMAL, nullptr);
    r.add(reinterpret_cast<void*>(&glx_square_staggered_gamma5_u1), OP_STAG_G5, nullptr);
    r.add(reinterpret_cast<void*>(&glx_gamma_5), OP_GAMMA5, nullptr);
    r.add(reinterpret_cast<void*>(&glx_square_staggered_deo_u1), OP_DEO, nullptr);
    r.add(reinterpret_cast<void*>(&glx_square_staggered_doe_u1), OP_DOE, nullptr);
    r.add(reinterpret_cast<void*>(&glx_square_staggered_m2mdeodoe_u1), OP_M2MDEODOE, nullptr);
    r.add(reinterpret_cast<void*>(&glx_apply_stencil_2d), OP_STENCIL, nullptr);
    r.add(reinterpret_cast<void*>(&glx_apply_stencil_2d_eo), OP_STENCIL_EO, nullptr);
    r.add(reinterpret_cast<void*>(&glx_apply_stencil_2d_oe), OP_STENCIL_OE, nullptr);
    r.add(reinterpret_cast<void*>(&glx_apply_square_staggered_m2mdeodoe_stencil), OP_STENCIL_M2MDEODOE, nullptr);
    r.add(reinterpret_cast<void*>(&glx_apply_stencil_2d_nc), OP_CST, nullptr);
    r.add(reinterpret_cast<void*>(&glx_apply_stencil_2d_nc_eo), OP_CST_EO, nullptr);
    r.add(reinterpret_cast<void*>(&glx_apply_stencil_2d_nc_oe), OP_CST_OE, nullptr);
    r.add(reinterpret_cast<void*>(&glx_apply_stencil_2d_nc_tb), OP_CST_TB, nullptr);
    r.add(reinterpret_cast<void*>(&glx_apply_stencil_2d_nc_bt), OP_CST_BT, nullptr);
    r.add(reinterpret_cast<void*>(&glx_apply_stencil_2d_nc_m2mdeodoe), OP_CST_M2MDEODOE, nullptr);
    r.add(reinterpret_cast<void*>(&glx_square_staggered), GLX_OP_STAGGERED_FREE, nullptr);
    r.add(reinterpret_cast<void*>(&glx_square_staggered_gamma5), GLX_OP_STAGGERED_GAMMA5_FREE, nullptr);
    r.add(reinterpret_cast<void*>(&glx_square_staggered_2linklaplace_u1), GLX_OP_STAGGERED_2LINKLAPLACE, nullptr);
    r.add(reinterpret_cast<void*>(&glx_staggered_symmshift_x), GLX_OP_SYMMSHIFT_X, nullptr);
    r.add(reinterpret_cast<void*>(&glx_staggered_symmshift_y), GLX_OP_SYMMSHIFT_Y, nullptr);
    r.add(reinterpret_cast<void*>(&glx_staggered_index_operator), GLX_OP_STAGGERED_INDEX, nullptr);
  }
};
BuiltinInit g_builtin_init;
}  // namespace

void glx::register_cst_callback(void* fn, int kind, int (*translate)(void*, glx_stencil_2d_nc*)) {
  if (!is_cst_kind(kind)) throw ArgError("register_cst_callback: not a general-stencil kind");
  Registry::get().add(fn, kind, nullptr, translate);
}

namespace {

// A device operator together with the device storage it points into.
struct Resolved {
  DevOperator op;
  DevBuf links, clover, hop, cs, cstmp;
};

// General stencil: checks the description, uploads the coefficients, and
// transposes them into the device layout (glx_cstencil.h).  Fills the
// operator's lattice, shifts and coefficient fields (not its kind).
void cst_upload(const glx_stencil_2d_nc& st, DevOperator& o, DevBuf& coef, DevBuf& tmp) {
  if (st.lx <= 0 || st.ly <= 0 || st.nc <= 0) throw ArgError("stencil dimensions and nc must be positive");
  if (double(st.lx) * st.ly * st.nc > 2147483647.0) throw ArgError("stencil vector length must fit an int");
  if (!st.clover || !st.hopping) throw ArgError("stencil needs clover and hopping coefficients");
  if (st.has_two && !st.two_link) throw ArgError("has_two is set but two_link is NULL");
  if (st.sdir < 0 || st.sdir > 13) throw ArgError("sdir must be a stencil_dir (0..13)");
  const int V = st.lx * st.ly, nterms = st.has_two ? 13 : 5;
  const size_t blk = size_t(V) * st.nc * st.nc * 16;
  {
    DevBuf stage(nterms * blk);
    char* d = static_cast<char*>(stage.get());
    GLX_HIP(hipMemcpy(d, st.clover, blk, hipMemcpyHostToDevice));
    GLX_HIP(hipMemcpy(d + blk, st.hopping, 4 * blk, hipMemcpyHostToDevice));
    if (st.has_two) GLX_HIP(hipMemcpy(d + 5 * blk, st.two_link, 8 * blk, hipMemcpyHostToDevice));
    coef = DevBuf(nterms * blk);
    hipLaunchKernelGGL(k_cstencil_transpose<0>, dim3(kMaxPointBlocks), dim3(kBlock), 0, nullptr, V, st.nc, nterms,
                       stage.as<c128>(), coef.as<c128>());
    GLX_CHECK_LAUNCH();
    GLX_HIP(hipDeviceSynchronize());
  }
  tmp = DevBuf(size_t(V) * st.nc * 16);
  o.lx = st.lx;
  o.ly = st.ly;
  o.nc = st.nc;
  o.shift = {st.shift[0], st.shift[1]};
  o.eo_shift = {st.eo_shift[0], st.eo_shift[1]};
  o.dof_shift = {st.dof_shift[0], st.dof_shift[1]};
  o.cs_coef = coef.as<c128>();
  o.cs_tmp = tmp.as<c128>();
  o.cs_sdir = st.sdir;
  o.cs_has_two = st.has_two != 0;
}

// The general-stencil description behind a registered callback's extra_info.
glx_stencil_2d_nc cst_desc(const RegEntry& e, void* extra) {
  glx_stencil_2d_nc st{};
  if (e.translate) {
    if (e.translate(extra, &st) != 0) throw ArgError("stencil description could not be read");
  } else {
    if (!extra) throw ArgError("general stencil operator needs a glx_stencil_2d_nc");
    st = *static_cast<const glx_stencil_2d_nc*>(extra);
  }
  return st;
}

size_t elem_bytes(bool cplx) { return cplx ? 16 : 8; }

void upload(DevBuf& buf, const void* host, size_t bytes) {
  buf = DevBuf(bytes);
  GLX_HIP(hipMemcpy(buf.get(), host, bytes, hipMemcpyHostToDevice));
}

bool is_built_kind(int k);
void build_op(int kind, const glx_staggered_u1_op& p, DevOperator& o, DevBuf& links, DevBuf& coef);

// Build the device operator for a solve or a direct application.
std::unique_ptr<Resolved> resolve(glx_matvec mv, void* extra, int size, bool cplx) {
  auto R = std::make_unique<Resolved>();
  RegEntry e;
  if (!Registry::get().find(reinterpret_cast<void*>(mv), e)) {
    if (!mv) throw ArgError("matrix_vector is NULL");
    R->op.host_cb = mv;
    R->op.host_extra = extra;
    R->op.host_complex = cplx;
    R->op.size_ = size;
    return R;
  }
  if (real_kind(e.kind) == cplx) throw ArgError("operator and vector types do not match (real vs complex)");
  DevOperator& o = R->op;
  o.kind = e.kind;
  if (is_cst_kind(e.kind)) {
    cst_upload(cst_desc(e, extra), o, R->cs, R->cstmp);
  } else if (is_built_kind(e.kind)) {
    const glx_staggered_u1_op* p = e.has_params ? &e.params : static_cast<const glx_staggered_u1_op*>(extra);
    if (!p) throw ArgError("operator needs extra_info (staggered_u1_op) or registered parameters");
    build_op(e.kind, *p, o, R->links, R->cs);
  } else if (is_stencil_kind(e.kind)) {
    const glx_stencil_2d* st = static_cast<const glx_stencil_2d*>(extra);
    if (!st || !st->hopping || !st->clover) throw ArgError("stencil operator needs a glx_stencil_2d with coefficients");
    o.lx = st->lx;
    o.ly = st->ly;
    o.nc = 1;
    const size_t V = size_t(o.lx) * o.ly;
    upload(R->clover, st->clover, V * 16);
    upload(R->hop, st->hopping, 4 * V * 16);
    o.clover = R->clover.as<c128>();
    o.hop = R->hop.as<c128>();
    o.shift = {st->shift[0], st->shift[1]};
    o.eo_shift = {st->eo_shift[0], st->eo_shift[1]};
    o.dof_shift = {st->dof_shift[0], st->dof_shift[1]};
  } else {
    const glx_staggered_u1_op* p = e.has_params ? &e.params : static_cast<const glx_staggered_u1_op*>(extra);
    if (!p) throw ArgError("operator needs extra_info (staggered_u1_op) or registered parameters");
    o.lx = p->x_fine;
    o.ly = p->y_fine;
    o.nc = (e.kind == OP_LAPLACE || e.kind == OP_LAPLACE_REAL) ? std::max(p->Nc, 1) : 1;
    o.mass = p->mass;
    if (op_links(e.kind)) {
      if (!p->lattice) throw ArgError("operator needs links (staggered_u1_op.lattice)");
      upload(R->links, p->lattice, size_t(o.lx) * o.ly * 2 * 16);
      o.u64 = R->links.as<c128>();
    }
  }
  if (o.lx <= 0 || o.ly <= 0) throw ArgError("lattice dimensions must be positive");
  if (size >= 0 && size != o.lx * o.ly * o.nc)
    throw ArgError("size " + std::to_string(size) + " does not match the operator's lattice (" +
                   std::to_string(o.lx * o.ly * o.nc) + ")");
  return R;
}

// Device copies of the caller's vectors.
struct DevVec {
  DevBuf buf;
  size_t bytes;
  DevVec(const double* host, size_t b) : bytes(b) {
    buf = DevBuf(b);
    if (host) GLX_HIP(hipMemcpy(buf.get(), host, b, hipMemcpyHostToDevice));
  }
  void to_host(double* host) const { GLX_HIP(hipMemcpy(host, buf.get(), bytes, hipMemcpyDeviceToHost)); }
  void* get() const { return buf.get(); }
};

void fill_info(glx_inversion_info* info, const SolveResult& r, const std::string& name) {
  if (!info) return;
  info->resSq = r.resSq;
  info->iter = r.iter;
  info->success = r.success ? 1 : 0;
  info->ops_count = r.ops_count;
  std::snprintf(info->name, sizeof(info->name), "%s", name.c_str());
  info->reliable_updates = r.reliable_updates;
}

enum SolverId { S_CG = 0, S_BICGSTAB = 1, S_GCR = 2 };
const char* solver_name(int s) { return s == S_CG ? "CG" : (s == S_BICGSTAB ? "BiCGStab" : "GCR"); }

// ops_count at the per-iteration print of iteration k (0-based), following
// where each reference solver increments it.
int ops_at_iter(int s, int k) {
  if (s == S_BICGSTAB) return 3 + 2 * k;
  return 2 + k;
}

SolveResult run_solver(int s, const DevOperator& op, void* x, const void* b, int max_iter, double eps, bool hist) {
  SolveOpts o;
  o.max_iter = max_iter;
  o.eps = eps;
  o.want_hist = hist;
  if (s == S_CG) return dev_cg(op, x, b, o);
  if (s == S_BICGSTAB) return dev_bicgstab(op, x, b, o);
  if (s == S_GCR) return dev_gcr(op, x, b, o);
  throw ArgError("unknown solver");
}

// One non-restarted solve with its verbosity lines.
SolveResult solve_and_print(int s, const DevOperator& op, void* x, const void* b, int max_iter, double eps,
                            glx_verbose* verb) {
  Printer pr(verb);
  const bool detail = pr.level() == GLX_VERB_DETAIL;
  SolveResult r = run_solver(s, op, x, b, max_iter, eps, detail);
  const std::string alg = solver_name(s);
  if (detail)
    for (int k = 0; k < int(r.relres_hist.size()); k++) pr.resid(alg, k + 1, ops_at_iter(s, k), r.relres_hist[k]);
  pr.summary(alg, r.success, r.iter, r.ops_count, std::sqrt(r.resSq) / r.bsqrt);
  return r;
}

// Restart driver: generic_cg.cpp:235-275 (and the BiCGStab / GCR copies).
SolveResult solve_restarted(int s, const DevOperator& op, void* x, const void* b, int max_iter, double eps,
                            int restart_freq, glx_verbose* verb, std::string& name) {
  std::ostringstream nm;
  nm << solver_name(s) << "(" << restart_freq << ")";
  name = nm.str();
  glx_verbose inner = restart_verbosity(verb);
  Printer pr(verb);
  int iter = 0, ops = 0;
  SolveResult r;
  do {
    if (verb) inner.out_len = verb->out_len;
    r = solve_and_print(s, op, x, b, restart_freq, eps, verb ? &inner : nullptr);
    if (verb) verb->out_len = inner.out_len;
    iter += r.iter;
    ops += r.ops_count;
    pr.restart(name, iter, ops, std::sqrt(r.resSq) / r.bsqrt);
  } while (iter < max_iter && !r.success && std::sqrt(r.resSq) / r.bsqrt > eps);
  r.iter = iter;
  r.ops_count = ops;
  pr.summary(name, r.success, iter, ops, std::sqrt(r.resSq) / r.bsqrt);
  r.success = !(std::sqrt(r.resSq) / r.bsqrt > eps);
  return r;
}

int host_solve(int s, bool cplx, double* phi, double* phi0, int size, int max_iter, double eps, int restart_freq,
               glx_matvec mv, void* extra, glx_verbose* verb, glx_inversion_info* info) {
  return guarded([&] {
    if (!phi || !phi0) throw ArgError("phi and phi0 must not be NULL");
    if (size <= 0) throw ArgError("size must be positive");
    auto R = resolve(mv, extra, size, cplx);
    const size_t bytes = size_t(size) * elem_bytes(cplx);
    DevVec x(phi, bytes), b(phi0, bytes);
    SolveResult r;
    std::string name = solver_name(s);
    if (restart_freq > 0) {
      r = solve_restarted(s, R->op, x.get(), b.get(), max_iter, eps, restart_freq, verb, name);
    } else {
      r = solve_and_print(s, R->op, x.get(), b.get(), max_iter, eps, verb);
    }
    x.to_host(phi);
    fill_info(info, r, name);
  });
}

int host_cg_m(bool cplx, double** phi, double* phi0, int n_shift, int size, int resid_freq_check, int max_iter,
              double eps, double* shifts, glx_matvec mv, void* extra, int worst_first, glx_verbose* verb,
              glx_inversion_info* info) {
  return guarded([&] {
    if (!phi || !phi0 || !shifts) throw ArgError("phi, phi0 and shifts must not be NULL");
    if (size <= 0 || n_shift <= 0) throw ArgError("size and n_shift must be positive");
    auto R = resolve(mv, extra, size, cplx);
    const size_t bytes = size_t(size) * elem_bytes(cplx);
    DevVec b(phi0, bytes);
    std::vector<std::unique_ptr<DevVec>> xs;
    std::vector<void*> xp;
    for (int k = 0; k < n_shift; k++) {
      xs.emplace_back(new DevVec(nullptr, bytes));
      xp.push_back(xs.back()->get());
    }
    Printer pr(verb);
    const bool detail = pr.level() == GLX_VERB_DETAIL;
    SolveOpts o;
    o.max_iter = max_iter;
    o.eps = eps;
    o.want_hist = detail;
    SolveResult r = dev_cg_m(R->op, xp.data(), b.get(), n_shift, shifts, resid_freq_check, worst_first != 0, o);
    for (int k = 0; k < n_shift; k++) xs[k]->to_host(phi[k]);
    if (detail)
      for (int k = 0; k < int(r.relres_hist.size()); k++) pr.resid("CG-M", k + 1, 1 + k, r.relres_hist[k]);
    std::vector<double> rel(n_shift);
    for (int k = 0; k < n_shift; k++) rel[k] = std::sqrt(r.resSqmrhs[k]) / r.bsqrt;
    pr.summary_multi(r.success, r.iter, r.ops_count, rel);
    fill_info(info, r, "CG-M");
    if (info && info->resSqmrhs)
      for (int k = 0; k < n_shift && (info->n_rhs <= 0 || k < info->n_rhs); k++) info->resSqmrhs[k] = r.resSqmrhs[k];
  });
}

// Direct application of a registered operator to host vectors.
void host_apply(glx_matvec self, double* lhs, double* rhs, void* extra, bool cplx) {
  int rc = guarded([&] {
    auto R = resolve(self, extra, -1, cplx);
    const size_t bytes = size_t(R->op.size()) * elem_bytes(cplx);
    DevVec in(rhs, bytes), out(nullptr, bytes);
    dev_apply(R->op, out.get(), in.get(), LINKS_F64, 0, nullptr);
    out.to_host(lhs);
  });
  if (rc != GLX_OK) {
    std::fprintf(stderr, "glx: operator application failed: %s\n", g_last_error.c_str());
    // the callback contract has no error channel: make the failure visible
    const glx_staggered_u1_op* p = static_cast<const glx_staggered_u1_op*>(extra);
    if (p && p->x_fine > 0 && p->y_fine > 0) {
      size_t n = size_t(p->x_fine) * p->y_fine * std::max(p->Nc, 1) * (cplx ? 2 : 1);
      for (size_t i = 0; i < n; i++) lhs[i] = std::nan("");
    }
  }
}

// ----------------------------------------------- stencil coefficient build
// operators_stencil.cpp:14-169: hopping blocks {+x, +y, -x, -y}; `sg` = +1
// for D, -1 for D^dag; g5 multiplies by eps(x) = (-1)^(x+y).
__global__ void k_build_stencil(int lx, int ly, const c128* __restrict__ u, c128* __restrict__ clover,
                                c128* __restrict__ hop, double sg, int g5) {
  const int V = lx * ly;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x) {
    const int x = i % lx, y = i / lx;
    const double eta = 1 - 2 * (x % 2);
    const double eo = g5 ? (((x + y) % 2 == 0) ? 1.0 : -1.0) : 1.0;
    const int xm = (x - 1 + lx) % lx, ym = (y - 1 + ly) % ly;
    const double h = -0.5 * sg * eo;
    clover[i] = {0.0, 0.0};
    hop[i] = h * u[2 * i];
    hop[V + i] = (h * eta) * u[2 * i + 1];
    hop[2 * V + i] = (-h) * conj(u[2 * (y * lx + xm)]);
    hop[3 * V + i] = (-h * eta) * conj(u[2 * (ym * lx + x) + 1]);
  }
}

int build_stencil(glx_stencil_2d* st, const glx_staggered_u1_op* op, double sg, int g5) {
  return guarded([&] {
    if (!st || !op || !op->lattice || !st->clover || !st->hopping) throw ArgError("NULL stencil or operator");
    const int lx = op->x_fine, ly = op->y_fine;
    if (lx <= 0 || ly <= 0) throw ArgError("lattice dimensions must be positive");
    const size_t V = size_t(lx) * ly;
    DevBuf u, cl(V * 16), hp(4 * V * 16);
    upload(u, op->lattice, 2 * V * 16);
    hipLaunchKernelGGL(k_build_stencil, dim3(point_blocks(int(V))), dim3(kBlock), 0, nullptr, lx, ly, u.as<c128>(),
                       cl.as<c128>(), hp.as<c128>(), sg, g5);
    GLX_CHECK_LAUNCH();
    GLX_HIP(hipMemcpy(st->clover, cl.get(), V * 16, hipMemcpyDeviceToHost));
    GLX_HIP(hipMemcpy(st->hopping, hp.get(), 4 * V * 16, hipMemcpyDeviceToHost));
    st->lx = lx;
    st->ly = ly;
    st->shift[0] = g5 ? 0.0 : op->mass;
    st->shift[1] = 0.0;
    st->eo_shift[0] = g5 ? op->mass : 0.0;
    st->eo_shift[1] = 0.0;
    st->dof_shift[0] = st->dof_shift[1] = 0.0;
  });
}

// The operators of operators.h off the solver path, as nc = 1 general
// stencils built on the device (13 terms, term-major: coef[t V + s]; for
// nc = 1 this is also the kernel's column-major layout).  Term order:
// clover, +x, +y, -x, -y, +2x, +x+y, +2y, -x+y, -2x, -x-y, -2y, +x-y.
enum BuiltOp : int {
  BUILT_STAG_FREE = GLX_OP_STAGGERED_FREE,        // operators.cpp:127-178
  BUILT_G5_FREE = GLX_OP_STAGGERED_GAMMA5_FREE,   // operators.cpp:262-313
  BUILT_2LINK = GLX_OP_STAGGERED_2LINKLAPLACE,    // operators.cpp:626-694
  BUILT_SYMM_X = GLX_OP_SYMMSHIFT_X,              // operators.cpp:697-736
  BUILT_SYMM_Y = GLX_OP_SYMMSHIFT_Y,              // operators.cpp:739-779
  BUILT_INDEX = GLX_OP_STAGGERED_INDEX,           // operators.cpp:782-846
};
bool is_built_kind(int k) { return k >= GLX_OP_STAGGERED_FREE && k <= GLX_OP_STAGGERED_INDEX; }
bool built_has_two(int k) { return k == BUILT_2LINK || k == BUILT_INDEX; }
bool built_needs_links(int k) { return k != BUILT_STAG_FREE && k != BUILT_G5_FREE; }

__global__ void k_build_op(int kind, int lx, int ly, const c128* __restrict__ u, double mass, double wc,
                           c128* __restrict__ coef) {
  const int V = lx * ly;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x) {
    const int x = i % lx, y = i / lx;
    const int xp = (x + 1) % lx, xm = (x - 1 + lx) % lx, yp = (y + 1) % ly, ym = (y - 1 + ly) % ly;
    auto site = [&](int a, int b) { return b * lx + a; };
    auto eta = [](int a) { return double(1 - 2 * (a % 2)); };
    const c128 z{0.0, 0.0};
    c128 t[kCsTerms];
#pragma unroll
    for (int k = 0; k < kCsTerms; k++) t[k] = z;
    const double e = eta(x);
    if (kind == BUILT_STAG_FREE || kind == BUILT_G5_FREE) {
      const double eo = kind == BUILT_G5_FREE ? (((x + y) % 2 == 0) ? 1.0 : -1.0) : 1.0;
      t[1] = {-0.5 * eo, 0.0};
      t[2] = {-0.5 * eo * e, 0.0};
      t[3] = {0.5 * eo, 0.0};
      t[4] = {0.5 * eo * e, 0.0};
    } else {
      const c128 ux = u[2 * i], uy = u[2 * i + 1];
      const c128 uxm = conj(u[2 * site(xm, y)]), uym = conj(u[2 * site(x, ym) + 1]);
      if (kind == BUILT_SYMM_X) {
        t[1] = 0.5 * ux;
        t[3] = 0.5 * uxm;
      } else if (kind == BUILT_SYMM_Y) {
        t[2] = (0.5 * e) * uy;
        t[4] = (0.5 * e) * uym;
      } else if (kind == BUILT_2LINK) {
        t[0] = {4.0 * wc, 0.0};
        t[1] = -0.5 * ux;
        t[2] = (-0.5 * e) * uy;
        t[3] = 0.5 * uxm;
        t[4] = (0.5 * e) * uym;
        const int xm2 = (x - 2 + 2 * lx) % lx, ym2 = (y - 2 + 2 * ly) % ly;
        t[5] = -wc * (ux * u[2 * site(xp, y)]);
        t[7] = -wc * (uy * u[2 * site(x, yp) + 1]);
        t[9] = -wc * (uxm * conj(u[2 * site(xm2, y)]));
        t[11] = -wc * (uym * conj(u[2 * site(x, ym2) + 1]));
      } else {  // BUILT_INDEX: i D(m = 0) - (i m / 2) (Gx Gy - Gy Gx), G the symmetric shifts
        const c128 I{0.0, 1.0};
        t[1] = I * (-0.5 * ux);
        t[2] = I * ((-0.5 * e) * uy);
        t[3] = I * (0.5 * uxm);
        t[4] = I * ((0.5 * e) * uym);
        // Gx Gy: the y shift acts at x +- 1 (its eta there), then the x shift
        const c128 uxp_y = u[2 * site(xp, y) + 1], uxm_y = u[2 * site(xm, y) + 1];
        const c128 uxp_ym = conj(u[2 * site(xp, ym) + 1]), uxm_ym = conj(u[2 * site(xm, ym) + 1]);
        const double exp_ = eta(xp), exm = eta(xm);
        const c128 a_pp = 0.25 * (ux * (exp_ * uxp_y)), a_pm = 0.25 * (ux * (exp_ * uxp_ym));
        const c128 a_mp = 0.25 * (uxm * (exm * uxm_y)), a_mm = 0.25 * (uxm * (exm * uxm_ym));
        // Gy Gx: the x shift acts at y +- 1, then the y shift (eta at x)
        const c128 uyp_x = u[2 * site(x, yp)], uyp_xm = conj(u[2 * site(xm, yp)]);
        const c128 uym_x = u[2 * site(x, ym)], uym_xm = conj(u[2 * site(xm, ym)]);
        const c128 b_pp = 0.25 * ((e * uy) * uyp_x), b_mp = 0.25 * ((e * uy) * uyp_xm);
        const c128 b_pm = 0.25 * ((e * uym) * uym_x), b_mm = 0.25 * ((e * uym) * uym_xm);
        const c128 f = c128{0.0, -0.5 * mass};  // -(i m / 2)
        t[6] = f * (a_pp - b_pp);   // +x+y
        t[8] = f * (a_mp - b_mp);   // -x+y
        t[10] = f * (a_mm - b_mm);  // -x-y
        t[12] = f * (a_pm - b_pm);  // +x-y
      }
    }
#pragma unroll
    for (int k = 0; k < kCsTerms; k++) coef[size_t(k) * V + i] = t[k];
  }
}

// Device operator for a built kind: links uploaded, coefficients built.
void build_op(int kind, const glx_staggered_u1_op& p, DevOperator& o, DevBuf& links, DevBuf& coef) {
  const int lx = p.x_fine, ly = p.y_fine;
  if (lx <= 0 || ly <= 0) throw ArgError("lattice dimensions must be positive");
  const size_t V = size_t(lx) * ly;
  if (built_needs_links(kind)) {
    if (!p.lattice) throw ArgError("operator needs links (staggered_u1_op.lattice)");
    upload(links, p.lattice, 2 * V * 16);
  }
  coef = DevBuf(size_t(kCsTerms) * V * 16);
  hipLaunchKernelGGL(k_build_op, dim3(point_blocks(int(V))), dim3(kBlock), 0, nullptr, kind, lx, ly,
                     built_needs_links(kind) ? links.as<c128>() : nullptr, p.mass, p.wilson_coeff, coef.as<c128>());
  GLX_CHECK_LAUNCH();
  GLX_HIP(hipDeviceSynchronize());
  o.kind = OP_CST;
  o.lx = lx;
  o.ly = ly;
  o.nc = 1;
  o.mass = p.mass;
  o.cs_coef = coef.as<c128>();
  o.cs_sdir = 0;
  o.cs_has_two = built_has_two(kind);
  o.shift = (kind == BUILT_STAG_FREE || kind == BUILT_2LINK) ? c128{p.mass, 0.0} : c128{0.0, 0.0};
  o.eo_shift = kind == BUILT_G5_FREE ? c128{p.mass, 0.0} : c128{0.0, 0.0};
  o.dof_shift = {0.0, 0.0};
}

// even/odd prepare and reconstruct on the device (operators.cpp:543-616,
// operators_stencil.cpp:179-236): one operator application plus a pointwise
// combination.
__global__ void k_eo_prepare(int lx, int ly, c128 m, const c128* __restrict__ b, c128* __restrict__ out) {
  const int V = lx * ly;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x) {
    const int x = i % lx, y = i / lx;
    if (((x + y) & 1) == 0) out[i] = m * b[i] - out[i];
  }
}
__global__ void k_eo_reconstruct(int lx, int ly, double inv_mass, const c128* __restrict__ xe,
                                 const c128* __restrict__ bo, c128* __restrict__ out) {
  const int V = lx * ly;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x) {
    const int x = i % lx, y = i / lx;
    if (((x + y) & 1) == 1) out[i] = inv_mass * (bo[i] - out[i]);
    else out[i] = xe[i];
  }
}

void eo_prepare(glx_matvec deo_fn, double* rhs_e, double* rhs, void* extra, bool stencil) {
  int rc = guarded([&] {
    auto R = resolve(deo_fn, extra, -1, true);
    const size_t V = size_t(R->op.lx) * R->op.ly;
    DevVec b(rhs, V * 16), out(nullptr, V * 16);
    dev_apply(R->op, out.get(), b.get(), LINKS_F64, 0, nullptr);
    c128 m = stencil ? R->op.shift : c128{R->op.mass, 0.0};
    hipLaunchKernelGGL(k_eo_prepare, dim3(point_blocks(int(V))), dim3(kBlock), 0, nullptr, R->op.lx, R->op.ly, m,
                       b.buf.as<c128>(), out.buf.as<c128>());
    GLX_CHECK_LAUNCH();
    out.to_host(rhs_e);
  });
  if (rc != GLX_OK) std::fprintf(stderr, "glx: eoprec_prepare failed: %s\n", g_last_error.c_str());
}
void eo_reconstruct(glx_matvec doe_fn, double* full, double* lhs_e, double* rhs_o, void* extra, bool stencil) {
  int rc = guarded([&] {
    auto R = resolve(doe_fn, extra, -1, true);
    const size_t V = size_t(R->op.lx) * R->op.ly;
    DevVec xe(lhs_e, V * 16), bo(rhs_o, V * 16), out(nullptr, V * 16);
    dev_apply(R->op, out.get(), xe.get(), LINKS_F64, 0, nullptr);
    const double inv_mass = 1.0 / (stencil ? R->op.shift.re : R->op.mass);
    hipLaunchKernelGGL(k_eo_reconstruct, dim3(point_blocks(int(V))), dim3(kBlock), 0, nullptr, R->op.lx, R->op.ly,
                       inv_mass, xe.buf.as<c128>(), bo.buf.as<c128>(), out.buf.as<c128>());
    GLX_CHECK_LAUNCH();
    out.to_host(full);
  });
  if (rc != GLX_OK) std::fprintf(stderr, "glx: eoprec_reconstruct failed: %s\n", g_last_error.c_str());
}

// General stencil through a callback: a two-link sdir on a stencil without
// two-link terms leaves lhs as it was (full / tb / bt), as the reference does.
void cs_host_apply(glx_matvec self, double* lhs, double* rhs, void* extra) {
  int rc = guarded([&] {
    RegEntry e;
    if (!Registry::get().find(reinterpret_cast<void*>(self), e)) throw ArgError("unregistered stencil callback");
    const glx_stencil_2d_nc st = cst_desc(e, extra);
    if (cst_piece(e.kind) >= 0 && cs_untouched(cst_piece(e.kind), st.sdir, st.has_two != 0)) return;
    auto R = resolve(self, extra, -1, true);
    const size_t bytes = size_t(R->op.size()) * 16;
    DevVec in(rhs, bytes), out(nullptr, bytes);
    dev_apply(R->op, out.get(), in.get(), LINKS_F64, 0, nullptr);
    out.to_host(lhs);
  });
  if (rc != GLX_OK) {
    std::fprintf(stderr, "glx: stencil application failed: %s\n", g_last_error.c_str());
    RegEntry e;
    if (Registry::get().find(reinterpret_cast<void*>(self), e)) {
      glx_stencil_2d_nc st{};
      bool ok = true;
      try {
        st = cst_desc(e, extra);
      } catch (...) {
        ok = false;
      }
      if (ok && st.lx > 0 && st.ly > 0 && st.nc > 0) {
        const size_t n = size_t(st.lx) * st.ly * st.nc * 2;
        for (size_t i = 0; i < n; i++) lhs[i] = std::nan("");
      }
    }
  }
}

// eo-prec prepare / reconstruct on a general stencil (operators_stencil.cpp:
// 179-192, 218-236): one stencil launch that also combines.
void cs_eo_helper(bool prepare, double* out_h, double* a_h, double* b_h, void* extra) {
  int rc = guarded([&] {
    auto R = resolve(reinterpret_cast<glx_matvec>(prepare ? &glx_apply_stencil_2d_nc_eo : &glx_apply_stencil_2d_nc_oe),
                     extra, -1, true);
    const size_t bytes = size_t(R->op.size()) * 16;
    DevVec a(a_h, bytes), out(nullptr, bytes);
    DevVec b(prepare ? nullptr : b_h, prepare ? 16 : bytes);
    CsDev d = cst_dev(R->op, prepare ? CS_EO : CS_OE);
    if (prepare) {
      d.comb_mode = CS_AXMY;
      d.comb_a = R->op.shift;
      d.comb = a.buf.as<c128>();
    } else {
      d.comb_mode = CS_RECON;
      d.comb_inv = 1.0 / R->op.shift.re;
      d.comb = b.buf.as<c128>();
      d.comb2 = a.buf.as<c128>();
    }
    launch_cst_one(nullptr, d, a.buf.as<c128>(), out.buf.as<c128>(), kNoGate);
    GLX_HIP(hipDeviceSynchronize());
    out.to_host(out_h);
  });
  if (rc != GLX_OK) std::fprintf(stderr, "glx: stencil eoprec helper failed: %s\n", g_last_error.c_str());
}

// generate_stencil_2d (coarse_stencil.cpp:1515-1627): the operator applied to
// each unit vector e_i, i = (site p, colour a); entry (q, c) of A e_i is the
// term-t coefficient (site q, row c, column a) where p = q + off_t.  The
// per-coefficient additions run in the reference's order.
void cs_generate(glx_stencil_2d_nc* st, glx_matvec mv, void* extra) {
  if (!st || !st->clover || !st->hopping || (st->has_two && !st->two_link)) throw ArgError("bad stencil arrays");
  if (!mv) throw ArgError("matrix_vector is NULL");
  const int lx = st->lx, ly = st->ly, nc = st->nc;
  if (lx <= 0 || ly <= 0 || nc <= 0) throw ArgError("stencil dimensions and nc must be positive");
  const int n = lx * ly * nc, nterms = st->has_two ? 13 : 5;
  const size_t blk = size_t(lx) * ly * nc * nc;
  std::vector<std::complex<double>> e(n), col(n);
  auto R = resolve(mv, extra, n, true);
  DevBuf de, dcol;
  if (!R->op.is_host()) {
    de = DevBuf(size_t(n) * 16);
    dcol = DevBuf(size_t(n) * 16);
    GLX_HIP(hipMemset(de.get(), 0, size_t(n) * 16));
  }
  const std::complex<double> one(1.0, 0.0), zero(0.0, 0.0);
  for (int i = 0; i < n; i++) {
    const int a = i % nc, p = i / nc, px = p % lx, py = p / lx;
    if (R->op.is_host()) {
      std::fill(e.begin(), e.end(), zero);
      e[i] = one;
      std::fill(col.begin(), col.end(), zero);
      mv(reinterpret_cast<double*>(col.data()), reinterpret_cast<double*>(e.data()), extra);
    } else {
      GLX_HIP(hipMemcpy(de.as<c128>() + i, &one, 16, hipMemcpyHostToDevice));
      dev_apply(R->op, dcol.get(), de.get(), LINKS_F64, 0, nullptr);
      GLX_HIP(hipMemcpy(col.data(), dcol.get(), size_t(n) * 16, hipMemcpyDeviceToHost));
      GLX_HIP(hipMemcpy(de.as<c128>() + i, &zero, 16, hipMemcpyHostToDevice));
    }
    for (int c = 0; c < nc; c++)
      for (int t = 0; t < nterms; t++) {
        int dx, dy;
        cs_offset(t, dx, dy);
        const int qx = ((px - dx) % lx + lx) % lx, qy = ((py - dy) % ly + ly) % ly;
        const size_t row = size_t(qy * lx + qx) * nc + c;
        double* base = t == 0 ? st->clover : (t < 5 ? st->hopping + 2 * (t - 1) * blk : st->two_link + 2 * (t - 5) * blk);
        std::complex<double>* coef = reinterpret_cast<std::complex<double>*>(base);
        coef[a + nc * row] += col[row];
      }
  }
}

}  // namespace

// =========================================================== lattices ====
struct glx_lattice_s {
  int lx = 0, ly = 0, flags = 0;
  DevBuf u64, u32, u16, sc16;
};

namespace {
// FP32 / FP16 link copies from the FP64 links (on the device).
__global__ void k_links_lowp(int n, const c128* __restrict__ u, c64* __restrict__ u32, HalfLink* __restrict__ u16,
                             float* __restrict__ sc) {
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const c128 a = u[2 * s], b = u[2 * s + 1];
    if (u32) {
      u32[2 * s] = {float(a.re), float(a.im)};
      u32[2 * s + 1] = {float(b.re), float(b.im)};
    }
    if (u16) {
      double m = fmax(fmax(fabs(a.re), fabs(a.im)), fmax(fabs(b.re), fabs(b.im)));
      if (m == 0.0) m = 1.0;
      const float scale = float(m);
      const double inv = 1.0 / double(scale);
      HalfLink h;
      h.x = __floats2half2_rn(float(a.re * inv), float(a.im * inv));
      h.y = __floats2half2_rn(float(b.re * inv), float(b.im * inv));
      u16[s] = h;
      sc[s] = scale;
    }
  }
}

// Refresh the FP32 / FP16 copies from the FP64 links.
void lattice_lowp(glx_lattice_s* L) {
  const size_t V = size_t(L->lx) * L->ly;
  if (L->flags & (GLX_LINKS_F32 | GLX_LINKS_F16)) {
    hipLaunchKernelGGL(k_links_lowp, dim3(point_blocks(int(V))), dim3(kBlock), 0, nullptr, int(V),
                       L->u64.as<c128>(), (L->flags & GLX_LINKS_F32) ? L->u32.as<c64>() : nullptr,
                       (L->flags & GLX_LINKS_F16) ? L->u16.as<HalfLink>() : nullptr,
                       (L->flags & GLX_LINKS_F16) ? L->sc16.as<float>() : nullptr);
    GLX_CHECK_LAUNCH();
  }
  GLX_HIP(hipDeviceSynchronize());
}

void lattice_fill(glx_lattice_s* L, const double* links) {
  const size_t V = size_t(L->lx) * L->ly;
  GLX_HIP(hipMemcpy(L->u64.get(), links, 2 * V * 16, hipMemcpyHostToDevice));
  lattice_lowp(L);
}

std::unique_ptr<glx_lattice_s> lattice_alloc(int lx, int ly, int flags) {
  if (lx <= 0 || ly <= 0) throw ArgError("lattice dimensions must be positive");
  if (double(lx) * ly > 2147483647.0) throw ArgError("lattice volume must fit a 32-bit site index");
  auto L = std::make_unique<glx_lattice_s>();
  L->lx = lx;
  L->ly = ly;
  L->flags = flags;
  const size_t V = size_t(lx) * ly;
  L->u64 = DevBuf(2 * V * 16);
  if (flags & GLX_LINKS_F32) L->u32 = DevBuf(2 * V * 8);
  if (flags & GLX_LINKS_F16) {
    L->u16 = DevBuf(V * sizeof(HalfLink));
    L->sc16 = DevBuf(V * sizeof(float));
  }
  return L;
}

DevOperator lattice_op(glx_lattice lat, int kind, double mass, int nc) {
  DevOperator o;
  o.kind = kind;
  o.mass = mass;
  o.nc = (kind == OP_LAPLACE || kind == OP_LAPLACE_REAL) ? std::max(nc, 1) : 1;
  if (is_stencil_kind(kind)) throw ArgError("stencil_2d operators take coefficients: use the host entry points");
  if (kind < 0 || (kind > OP_M2MDEODOE && kind != OP_STAG_EO_SCHUR)) throw ArgError("unknown operator kind");
  if (op_links(kind) && !lat) throw ArgError("this operator needs a lattice with links");
  if (lat) {
    o.lx = lat->lx;
    o.ly = lat->ly;
    o.u64 = lat->u64.as<c128>();
    o.u32 = (lat->flags & GLX_LINKS_F32) ? lat->u32.as<c64>() : nullptr;
    o.u16 = (lat->flags & GLX_LINKS_F16) ? lat->u16.as<HalfLink>() : nullptr;
    o.sc16 = (lat->flags & GLX_LINKS_F16) ? lat->sc16.as<float>() : nullptr;
  }
  return o;
}
}  // namespace

namespace {
template <class F, class FR>
int dispatch_unpreconditioned(F plain, FR restarted, double* lhs, double* rhs, int size, int type,
                              const glx_inverter_params* p, glx_matvec mv, void* extra, glx_verbose* verb,
                              glx_inversion_info* info) {
  if (!p) {
    g_last_error = "params is NULL";
    return GLX_ERR_ARG;
  }
  if (type == -1) {  // MINV_INVALID: the reference returns a default inversion_info
    if (info) {
      std::memset(info, 0, sizeof(*info));
      info->n_rhs = -1;
    }
    return GLX_OK;
  }
  if (type != 0 && type != 2 && type != 3) {
    g_last_error = "minv_unpreconditioned: solver " + std::to_string(type) +
                   " is not part of this library (CG 0, GCR 2 and BiCGStab 3 are)";
    return GLX_ERR_ARG;
  }
  const int k = type == 0 ? 0 : (type == 2 ? 2 : 1);  // index into {CG, BiCGStab, GCR}
  return p->restart ? restarted[k](lhs, rhs, size, p->max_iters, p->tol, p->restart_freq, mv, extra, verb, info)
                    : plain[k](lhs, rhs, size, p->max_iters, p->tol, mv, extra, verb, info);
}
}  // namespace

namespace {
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_stream(const float4* __restrict__ a, const float4* __restrict__ b,
                                                    const float4* __restrict__ c, float4* __restrict__ o, long n,
                                                    float4* __restrict__ sink) {
  const long stride = long(gridDim.x) * kBlock;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  for (long i = long(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) {
    if constexpr (MODE == 0) {
      const float4 v = a[i];
      acc.x += v.x;
      acc.y += v.y;
      acc.z += v.z;
      acc.w += v.w;
    } else if constexpr (MODE == 1) {
      o[i] = make_float4(1.f, 2.f, 3.f, float(i));
    } else if constexpr (MODE == 2) {
      o[i] = a[i];
    } else {
      const float4 x = a[i], y = b[i], z = c[i];
      o[i] = make_float4(x.x + y.x + z.x, x.y + y.y + z.y, x.z + y.z + z.z, x.w + y.w + z.w);
    }
  }
  // keeps the reads of the read-only pass alive; never true for finite data
  if constexpr (MODE == 0) {
    if (acc.x == -1.2345e30f) sink[threadIdx.x] = acc;
  }
}
}  // namespace

// ====================================================== exported C ABI ====
extern "C" {

void glx_square_laplace_real(double* lhs, double* rhs, void* extra) { host_apply(&glx_square_laplace_real, lhs, rhs, extra, false); }
void glx_square_laplace(double* lhs, double* rhs, void* extra) { host_apply(&glx_square_laplace, lhs, rhs, extra, true); }
void glx_square_laplace_u1(double* lhs, double* rhs, void* extra) { host_apply(&glx_square_laplace_u1, lhs, rhs, extra, true); }
void glx_square_staggered_u1(double* lhs, double* rhs, void* extra) { host_apply(&glx_square_staggered_u1, lhs, rhs, extra, true); }
void glx_gamma_5(double* lhs, double* rhs, void* extra) { host_apply(&glx_gamma_5, lhs, rhs, extra, true); }
void glx_square_staggered_gamma5_u1(double* lhs, double* rhs, void* extra) { host_apply(&glx_square_staggered_gamma5_u1, lhs, rhs, extra, true); }
void glx_square_staggered_dagger_u1(double* lhs, double* rhs, void* extra) { host_apply(&glx_square_staggered_dagger_u1, lhs, rhs, extra, true); }
void glx_square_staggered_normal_u1(double* lhs, double* rhs, void* extra) { host_apply(&glx_square_staggered_normal_u1, lhs, rhs, extra, true); }
void glx_square_staggered_deo_u1(double* lhs, double* rhs, void* extra) { host_apply(&glx_square_staggered_deo_u1, lhs, rhs, extra, true); }
void glx_square_staggered_doe_u1(double* lhs, double* rhs, void* extra) { host_apply(&glx_square_staggered_doe_u1, lhs, rhs, extra, true); }
void glx_square_staggered_m2mdeodoe_u1(double* lhs, double* rhs, void* extra) { host_apply(&glx_square_staggered_m2mdeodoe_u1, lhs, rhs, extra, true); }
void glx_square_staggered_eoprec_prepare(double* rhs_e, double* rhs_orig, void* extra) {
  eo_prepare(&glx_square_staggered_deo_u1, rhs_e, rhs_orig, extra, false);
}
void glx_square_staggered_eoprec_reconstruct(double* lhs_full, double* lhs_e, double* rhs_o, void* extra) {
  eo_reconstruct(&glx_square_staggered_doe_u1, lhs_full, lhs_e, rhs_o, extra, false);
}

int glx_get_square_staggered_u1_stencil(glx_stencil_2d* st, const glx_staggered_u1_op* op) { return build_stencil(st, op, +1.0, 0); }
int glx_get_square_staggered_gamma5_u1_stencil(glx_stencil_2d* st, const glx_staggered_u1_op* op) { return build_stencil(st, op, +1.0, 1); }
int glx_get_square_staggered_dagger_u1_stencil(glx_stencil_2d* st, const glx_staggered_u1_op* op) { return build_stencil(st, op, -1.0, 0); }
void glx_apply_stencil_2d(double* lhs, double* rhs, void* st) { host_apply(&glx_apply_stencil_2d, lhs, rhs, st, true); }
void glx_apply_stencil_2d_eo(double* lhs, double* rhs, void* st) { host_apply(&glx_apply_stencil_2d_eo, lhs, rhs, st, true); }
void glx_apply_stencil_2d_oe(double* lhs, double* rhs, void* st) { host_apply(&glx_apply_stencil_2d_oe, lhs, rhs, st, true); }
void glx_apply_stencil_2d_nc(double* lhs, double* rhs, void* st) { cs_host_apply(&glx_apply_stencil_2d_nc, lhs, rhs, st); }
void glx_square_staggered(double* lhs, double* rhs, void* e) { host_apply(&glx_square_staggered, lhs, rhs, e, true); }
void glx_square_staggered_gamma5(double* lhs, double* rhs, void* e) {
  host_apply(&glx_square_staggered_gamma5, lhs, rhs, e, true);
}
void glx_square_staggered_2linklaplace_u1(double* lhs, double* rhs, void* e) {
  host_apply(&glx_square_staggered_2linklaplace_u1, lhs, rhs, e, true);
}
void glx_staggered_symmshift_x(double* lhs, double* rhs, void* e) {
  host_apply(&glx_staggered_symmshift_x, lhs, rhs, e, true);
}
void glx_staggered_symmshift_y(double* lhs, double* rhs, void* e) {
  host_apply(&glx_staggered_symmshift_y, lhs, rhs, e, true);
}
void glx_staggered_index_operator(double* lhs, double* rhs, void* e) {
  host_apply(&glx_staggered_index_operator, lhs, rhs, e, true);
}
void glx_apply_stencil_2d_nc_eo(double* lhs, double* rhs, void* st) {
  cs_host_apply(&glx_apply_stencil_2d_nc_eo, lhs, rhs, st);
}
void glx_apply_stencil_2d_nc_oe(double* lhs, double* rhs, void* st) {
  cs_host_apply(&glx_apply_stencil_2d_nc_oe, lhs, rhs, st);
}
void glx_apply_stencil_2d_nc_tb(double* lhs, double* rhs, void* st) {
  cs_host_apply(&glx_apply_stencil_2d_nc_tb, lhs, rhs, st);
}
void glx_apply_stencil_2d_nc_bt(double* lhs, double* rhs, void* st) {
  cs_host_apply(&glx_apply_stencil_2d_nc_bt, lhs, rhs, st);
}
void glx_apply_stencil_2d_nc_m2mdeodoe(double* lhs, double* rhs, void* st) {
  cs_host_apply(&glx_apply_stencil_2d_nc_m2mdeodoe, lhs, rhs, st);
}
void glx_stencil_2d_nc_eoprec_prepare(double* rhs_e, double* rhs_orig, void* st) {
  cs_eo_helper(true, rhs_e, rhs_orig, nullptr, st);
}
void glx_stencil_2d_nc_eoprec_reconstruct(double* lhs_full, double* lhs_e, double* rhs_o, void* st) {
  cs_eo_helper(false, lhs_full, lhs_e, rhs_o, st);
}
int glx_generate_stencil_2d_nc(glx_stencil_2d_nc* st, glx_matvec mv, void* extra) {
  return guarded([&] { cs_generate(st, mv, extra); });
}
void glx_apply_square_staggered_m2mdeodoe_stencil(double* lhs, double* rhs, void* st) {
  host_apply(&glx_apply_square_staggered_m2mdeodoe_stencil, lhs, rhs, st, true);
}
void glx_apply_square_staggered_eoprec_prepare_stencil(double* rhs_e, double* rhs_orig, void* st) {
  eo_prepare(&glx_apply_stencil_2d_eo, rhs_e, rhs_orig, st, true);
}
void glx_apply_square_staggered_eoprec_reconstruct_stencil(double* lhs_full, double* lhs_e, double* rhs_o, void* st) {
  eo_reconstruct(&glx_apply_stencil_2d_oe, lhs_full, lhs_e, rhs_o, st, true);
}

int glx_register_operator(glx_matvec callback, int kind, const glx_staggered_u1_op* params) {
  return guarded([&] {
    if (!callback) throw ArgError("callback is NULL");
    if ((kind < 0 || kind > OP_STENCIL_M2MDEODOE) && !is_built_kind(kind)) throw ArgError("unknown operator kind");
    Registry::get().add(reinterpret_cast<void*>(callback), kind, params);
  });
}
int glx_unregister_operator(glx_matvec callback) {
  return guarded([&] { Registry::get().remove(reinterpret_cast<void*>(callback)); });
}

#define GLX_SOLVE_ENTRY(NAME, S, CPLX)                                                                         \
  int NAME(double* phi, double* phi0, int size, int max_iter, double eps, glx_matvec mv, void* extra,            \
           glx_verbose* verb, glx_inversion_info* info) {                                                      \
    return host_solve(S, CPLX, phi, phi0, size, max_iter, eps, 0, mv, extra, verb, info);                      \
  }
#define GLX_RESTART_ENTRY(NAME, S, CPLX)                                                                       \
  int NAME(double* phi, double* phi0, int size, int max_iter, double eps, int restart_freq, glx_matvec mv,       \
           void* extra, glx_verbose* verb, glx_inversion_info* info) {                                         \
    if (restart_freq <= 0) {                                                                                   \
      g_last_error = "restart_freq must be positive";                                                          \
      return GLX_ERR_ARG;                                                                                      \
    }                                                                                                          \
    return host_solve(S, CPLX, phi, phi0, size, max_iter, eps, restart_freq, mv, extra, verb, info);           \
  }
GLX_SOLVE_ENTRY(glx_minv_vector_cg_r, S_CG, false)
GLX_SOLVE_ENTRY(glx_minv_vector_cg_c, S_CG, true)
GLX_SOLVE_ENTRY(glx_minv_vector_bicgstab_r, S_BICGSTAB, false)
GLX_SOLVE_ENTRY(glx_minv_vector_bicgstab_c, S_BICGSTAB, true)
GLX_SOLVE_ENTRY(glx_minv_vector_gcr_r, S_GCR, false)
GLX_SOLVE_ENTRY(glx_minv_vector_gcr_c, S_GCR, true)
GLX_RESTART_ENTRY(glx_minv_vector_cg_restart_r, S_CG, false)
GLX_RESTART_ENTRY(glx_minv_vector_cg_restart_c, S_CG, true)
GLX_RESTART_ENTRY(glx_minv_vector_bicgstab_restart_r, S_BICGSTAB, false)
GLX_RESTART_ENTRY(glx_minv_vector_bicgstab_restart_c, S_BICGSTAB, true)
GLX_RESTART_ENTRY(glx_minv_vector_gcr_restart_r, S_GCR, false)
GLX_RESTART_ENTRY(glx_minv_vector_gcr_restart_c, S_GCR, true)
#undef GLX_SOLVE_ENTRY
#undef GLX_RESTART_ENTRY

int glx_minv_vector_cg_m_r(double** phi, double* phi0, int n_shift, int size, int resid_freq_check, int max_iter,
                           double eps, double* shifts, glx_matvec mv, void* extra, int worst_first,
                           glx_verbose* verb, glx_inversion_info* info) {
  return host_cg_m(false, phi, phi0, n_shift, size, resid_freq_check, max_iter, eps, shifts, mv, extra, worst_first,
                   verb, info);
}
int glx_minv_vector_cg_m_c(double** phi, double* phi0, int n_shift, int size, int resid_freq_check, int max_iter,
                           double eps, double* shifts, glx_matvec mv, void* extra, int worst_first,
                           glx_verbose* verb, glx_inversion_info* info) {
  return host_cg_m(true, phi, phi0, n_shift, size, resid_freq_check, max_iter, eps, shifts, mv, extra, worst_first,
                   verb, info);
}

// ---------------------------------------------------------- device API --
int glx_lattice_create(glx_lattice* lat, const double* links, int lx, int ly, int flags) {
  return guarded([&] {
    if (!lat || !links) throw ArgError("lat and links must not be NULL");
    auto L = lattice_alloc(lx, ly, flags);
    lattice_fill(L.get(), links);
    *lat = L.release();
  });
}
int glx_lattice_create_gaussian(glx_lattice* lat, int lx, int ly, int flags, double beta, unsigned long long seed,
                                unsigned int stream) {
  return guarded([&] {
    if (!lat) throw ArgError("lat must not be NULL");
    if (!std::isfinite(beta)) throw ArgError("beta must be finite");
    auto L = lattice_alloc(lx, ly, flags);
    dev_gauss_links(L->u64.get(), lx, ly, beta, seed, stream, nullptr);
    lattice_lowp(L.get());
    *lat = L.release();
  });
}
int glx_lattice_gaussian_links(glx_lattice lat, double beta, unsigned long long seed, unsigned int stream) {
  return guarded([&] {
    if (!lat) throw ArgError("lat must not be NULL");
    if (!std::isfinite(beta)) throw ArgError("beta must be finite");
    GLX_HIP(hipDeviceSynchronize());
    dev_gauss_links(lat->u64.get(), lat->lx, lat->ly, beta, seed, stream, nullptr);
    lattice_lowp(lat);
  });
}
int glx_lattice_get_links(glx_lattice lat, double* links) {
  return guarded([&] {
    if (!lat || !links) throw ArgError("lat and links must not be NULL");
    GLX_HIP(hipDeviceSynchronize());
    GLX_HIP(hipMemcpy(links, lat->u64.get(), size_t(lat->lx) * lat->ly * 32, hipMemcpyDeviceToHost));
  });
}
int glx_lattice_plaquette(glx_lattice lat, double* plaq) {
  return guarded([&] {
    if (!lat || !plaq) throw ArgError("lat and plaq must not be NULL");
    GLX_HIP(hipDeviceSynchronize());
    const std::complex<double> p = dev_plaquette(lat->u64.get(), lat->lx, lat->ly, nullptr);
    plaq[0] = p.real();
    plaq[1] = p.imag();
  });
}
int glx_lattice_update_links(glx_lattice lat, const double* links) {
  return guarded([&] {
    if (!lat || !links) throw ArgError("lat and links must not be NULL");
    lattice_fill(lat, links);
  });
}
int glx_lattice_destroy(glx_lattice lat) {
  return guarded([&] {
    GLX_HIP(hipDeviceSynchronize());
    delete lat;
  });
}

int glx_dev_solve(int solver, glx_lattice lat, int kind, double mass, int nc, double* x_dev, const double* b_dev,
                  int max_iter, double eps, glx_verbose* verb, glx_inversion_info* info) {
  return guarded([&] {
    if (!x_dev || !b_dev) throw ArgError("x_dev and b_dev must not be NULL");
    DevOperator o = lattice_op(lat, kind, mass, nc);
    if (o.lx <= 0) throw ArgError("operator without links needs a lattice for its dimensions");
    SolveResult r = solve_and_print(solver, o, x_dev, b_dev, max_iter, eps, verb);
    fill_info(info, r, solver_name(solver));
  });
}

int glx_dev_cg_m(glx_lattice lat, int kind, double mass, int nc, double** x_dev, const double* b_dev, int n_shift,
                 const double* shifts, int resid_freq_check, int max_iter, double eps, int worst_first,
                 glx_verbose* verb, glx_inversion_info* info) {
  return guarded([&] {
    if (!x_dev || !b_dev || !shifts) throw ArgError("x_dev, b_dev and shifts must not be NULL");
    DevOperator o = lattice_op(lat, kind, mass, nc);
    if (o.lx <= 0) throw ArgError("operator without links needs a lattice for its dimensions");
    Printer pr(verb);
    SolveOpts so;
    so.max_iter = max_iter;
    so.eps = eps;
    so.want_hist = pr.level() == GLX_VERB_DETAIL;
    std::vector<void*> xp(x_dev, x_dev + n_shift);
    SolveResult r = dev_cg_m(o, xp.data(), b_dev, n_shift, shifts, resid_freq_check, worst_first != 0, so);
    for (int k = 0; k < int(r.relres_hist.size()); k++) pr.resid("CG-M", k + 1, 1 + k, r.relres_hist[k]);
    std::vector<double> rel(n_shift);
    for (int k = 0; k < n_shift; k++) rel[k] = std::sqrt(r.resSqmrhs[k]) / r.bsqrt;
    pr.summary_multi(r.success, r.iter, r.ops_count, rel);
    fill_info(info, r, "CG-M");
    if (info && info->resSqmrhs)
      for (int k = 0; k < n_shift && (info->n_rhs <= 0 || k < info->n_rhs); k++) info->resSqmrhs[k] = r.resSqmrhs[k];
  });
}

int glx_dev_cg_m_batched(int n_sys, const glx_batch_system* sys, int n_shift, const double* shifts,
                         int resid_freq_check, int max_iter, double eps, int worst_first, int use_graph,
                         glx_inversion_info* infos) {
  return glx_dev_cg_m_batched_ex(n_sys, sys, n_shift, shifts, resid_freq_check, max_iter, eps, worst_first, use_graph,
                                 0, infos, nullptr);
}

int glx_dev_cg_m_batched_ex(int n_sys, const glx_batch_system* sys, int n_shift, const double* shifts,
                            int resid_freq_check, int max_iter, double eps, int worst_first, int use_graph,
                            int time_phases, glx_inversion_info* infos, glx_batch_stats* stats) {
  return guarded([&] {
    if (n_sys < 0 || (n_sys > 0 && !sys) || !shifts) throw ArgError("sys and shifts must not be NULL");
    std::vector<BatchSystem> bs(n_sys);
    std::vector<std::vector<void*>> xp(n_sys);
    for (int s = 0; s < n_sys; s++) {
      if (!sys[s].lat || !sys[s].b_dev || !sys[s].x_dev) throw ArgError("batched system with a NULL member");
      bs[s].op = lattice_op(sys[s].lat, OP_STAG_NORMAL, sys[s].mass, 1);
      bs[s].b = sys[s].b_dev;
      xp[s].assign(sys[s].x_dev, sys[s].x_dev + n_shift);
      bs[s].xs = xp[s].data();
    }
    SolveOpts so;
    so.max_iter = max_iter;
    so.eps = eps;
    so.group = g_group;
    so.use_graph = use_graph != 0;
    BatchStats bst;
    std::vector<SolveResult> r = dev_cg_m_batched(bs, n_shift, shifts, resid_freq_check, worst_first != 0, so,
                                                  stats ? &bst : nullptr, time_phases != 0);
    if (stats) {
      stats->bytes_stencil = bst.bytes_stencil;
      stats->bytes_r_pass = bst.bytes_r;
      stats->bytes_shift_pass = bst.bytes_shift;
      stats->ms_stencil = bst.ms_stencil;
      stats->ms_r_pass = bst.ms_r;
      stats->ms_shift_pass = bst.ms_shift;
      stats->iterations = bst.iterations;
      stats->shift_updates_x = bst.upd_x;
      stats->shift_updates_p = bst.upd_p;
    }
    if (infos) {
      for (int s = 0; s < n_sys; s++) {
        fill_info(&infos[s], r[s], "CG-M");
        if (infos[s].resSqmrhs)
          for (int k = 0; k < n_shift && (infos[s].n_rhs <= 0 || k < infos[s].n_rhs); k++)
            infos[s].resSqmrhs[k] = r[s].resSqmrhs[k];
      }
    }
  });
}

int glx_dev_cg_mixed(glx_lattice lat, int kind, double mass, double* x_dev, const double* b_dev, int max_iter,
                     double eps, double delta, int link_prec, glx_verbose* verb, glx_inversion_info* info) {
  return guarded([&] {
    if (!x_dev || !b_dev) throw ArgError("x_dev and b_dev must not be NULL");
    if (!(delta > 0.0 && delta < 1.0)) throw ArgError("delta must be in (0, 1)");
    DevOperator o = lattice_op(lat, kind, mass, 1);
    if (o.lx <= 0) throw ArgError("operator without links needs a lattice for its dimensions");
    if (op_links(kind)) {
      if (link_prec == LINKS_F16 && !o.u16) throw ArgError("lattice was created without GLX_LINKS_F16");
      if (link_prec != LINKS_F16 && !o.u32) throw ArgError("lattice was created without GLX_LINKS_F32");
    }
    Printer pr(verb);
    SolveOpts so;
    so.max_iter = max_iter;
    so.eps = eps;
    so.want_hist = pr.level() == GLX_VERB_DETAIL;
    so.group = g_group;
    so.use_graph = g_use_graph != 0;
    MixedOpts mo;
    mo.link_prec = link_prec == LINKS_F16 ? LINKS_F16 : LINKS_F32;
    mo.delta = delta;
    SolveResult r = dev_cg_mixed(o, x_dev, b_dev, so, mo);
    for (int k = 0; k < int(r.relres_hist.size()); k++) pr.resid("CG-mixed", k + 1, 2 + k, r.relres_hist[k]);
    pr.summary("CG-mixed", r.success, r.iter, r.ops_count, std::sqrt(r.resSq) / r.bsqrt);
    fill_info(info, r, "CG-mixed");
  });
}

// ----------------------------------------------- allocation-free mixed CG --
// The mixed-precision CG above, with every device buffer of the solve (state,
// reduction partials, residual history, the FP32 Krylov basis or vectors, the
// FP64 reliable-update vector) carved out of `scratch`, one per-GPU array of
// kScratchBytes held for the life of the process.  The solve runs on the
// calling thread's per-thread default stream and creates no stream, event,
// graph or pinned host memory; it polls its state synchronously after each
// group of iterations.  Calls on one GPU share its array: calls that overlap
// on one GPU (other threads) race.  Sizes: at most kNoAllocSites lattice
// sites (the 104 MiB of the s-step form at 1024 x 1024 fit) and
// kNoAllocMaxIter iterations (the history buffer).
namespace {
constexpr size_t kScratchBytes = size_t(112) << 20;
constexpr long kNoAllocSites = 1L << 20;
constexpr int kNoAllocMaxIter = 1 << 18;
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) char scratch[kScratchBytes];

struct ScratchAddr {
  std::once_flag once;
  void* ptr = nullptr;
  hipError_t err = hipSuccess;
};
ScratchAddr g_scratch_addr[kNoAllocMaxDevices];

// The device address of this device's `scratch`, looked up once per device
// with that device current.  Host code uses only this address: the array's
// own name is a host-side stand-in.
void* scratch_on_current_device() {
  int dev = 0;
  GLX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) throw ArgError("device index beyond the allocation-free table");
  ScratchAddr& a = g_scratch_addr[dev];
  std::call_once(a.once, [&] { a.err = hipGetSymbolAddress(&a.ptr, HIP_SYMBOL(scratch)); });
  if (a.err != hipSuccess || !a.ptr)
    throw HipError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(a.err));
  return a.ptr;
}
}  // namespace

int glx_dev_cg_mixed_noalloc(glx_lattice lat, int kind, double mass, double* x_dev, const double* b_dev,
                             int max_iter, double eps, double delta, int link_prec, glx_verbose* verb,
                             glx_inversion_info* info) {
  return guarded([&] {
    if (!x_dev || !b_dev) throw ArgError("x_dev and b_dev must not be NULL");
    if (!(delta > 0.0 && delta < 1.0)) throw ArgError("delta must be in (0, 1)");
    DevOperator o = lattice_op(lat, kind, mass, 1);
    if (o.lx <= 0) throw ArgError("operator without links needs a lattice for its dimensions");
    if (op_links(kind)) {
      if (link_prec == LINKS_F16 && !o.u16) throw ArgError("lattice was created without GLX_LINKS_F16");
      if (link_prec != LINKS_F16 && !o.u32) throw ArgError("lattice was created without GLX_LINKS_F32");
    }
    if (long(o.size()) > kNoAllocSites)
      throw ArgError("glx_dev_cg_mixed_noalloc: " + std::to_string(o.size()) + " sites exceed its maximum of " +
                     std::to_string(kNoAllocSites) + " (1024 x 1024); use glx_dev_cg_mixed");
    if (max_iter > kNoAllocMaxIter)
      throw ArgError("glx_dev_cg_mixed_noalloc: max_iter beyond its maximum of " + std::to_string(kNoAllocMaxIter));
    void* base = scratch_on_current_device();
    Printer pr(verb);
    SolveOpts so;
    so.max_iter = max_iter;
    so.eps = eps;
    so.want_hist = pr.level() == GLX_VERB_DETAIL;
    so.group = g_group;
    so.use_graph = false;
    so.stream = hipStreamPerThread;
    MixedOpts mo;
    mo.link_prec = link_prec == LINKS_F16 ? LINKS_F16 : LINKS_F32;
    mo.delta = delta;
    SolveResult r;
    {
      NoAllocScope scope(base, kScratchBytes, hipStreamPerThread);
      r = dev_cg_mixed(o, x_dev, b_dev, so, mo);
    }
    for (int k = 0; k < int(r.relres_hist.size()); k++) pr.resid("CG-mixed", k + 1, 2 + k, r.relres_hist[k]);
    pr.summary("CG-mixed", r.success, r.iter, r.ops_count, std::sqrt(r.resSq) / r.bsqrt);
    fill_info(info, r, "CG-mixed");
  });
}

int glx_dev_staggered_grad(glx_lattice lat, double mass, const double* phi_dev, int precision, int max_iter,
                           double eps, double* grad_dev, double* chi_dev, double* dS_dm, double* S, double* S_robust,
                           glx_inversion_info* info) {
  return guarded([&] {
    if (!phi_dev || !grad_dev) throw ArgError("phi_dev and grad_dev must not be NULL");
    if (precision < 0 || precision > 2) throw ArgError("precision must be 0 (FP64), 1 (FP32 links) or 2 (FP16 links)");
    DevOperator normal = lattice_op(lat, OP_STAG_NORMAL, mass, 1);
    DevOperator dop = lattice_op(lat, OP_STAG, mass, 1);
    if (precision == LINKS_F16 && !normal.u16) throw ArgError("lattice was created without GLX_LINKS_F16");
    if (precision == LINKS_F32 && !normal.u32) throw ArgError("lattice was created without GLX_LINKS_F32");
    SolveOpts so;
    so.max_iter = max_iter;
    so.eps = eps;
    so.group = g_group;
    so.use_graph = g_use_graph != 0;
    GradResult g = dev_staggered_grad(normal, dop, phi_dev, precision, so, grad_dev, chi_dev);
    if (dS_dm) *dS_dm = g.dS_dm;
    if (S) *S = g.S;
    if (S_robust) *S_robust = g.S_robust;
    fill_info(info, g.solve, g.solve.name.c_str());
  });
}

int glx_dev_solve_staggered_eo(glx_lattice lat, double mass, double* x_dev, const double* b_dev, int method,
                               int max_iter, double eps, double* full_resSq, glx_inversion_info* info) {
  return guarded([&] {
    if (!x_dev || !b_dev) throw ArgError("x_dev and b_dev must not be NULL");
    if (method != 0 && method != 1 && method != 2 && method != 3) throw ArgError("method must be 0, 1, 2 or 3");
    DevOperator o = lattice_op(lat, OP_STAG, mass, 1);
    if (method == LINKS_F16 && !o.u16) throw ArgError("lattice was created without GLX_LINKS_F16");
    if (method == LINKS_F32 && !o.u32) throw ArgError("lattice was created without GLX_LINKS_F32");
    SolveOpts so;
    so.max_iter = max_iter;
    so.eps = eps;
    so.group = g_group;
    so.use_graph = g_use_graph != 0;
    SolveResult r = dev_solve_staggered_eo(o, x_dev, b_dev, method, so, full_resSq);
    fill_info(info, r, r.name);
  });
}
int glx_dev_eo_prepare(glx_lattice lat, double mass, double* be_dev, const double* b_dev) {
  return guarded([&] {
    if (!be_dev || !b_dev) throw ArgError("be_dev and b_dev must not be NULL");
    dev_eo_prepare(lattice_op(lat, OP_STAG, mass, 1), be_dev, b_dev, nullptr);
  });
}
int glx_dev_eo_reconstruct(glx_lattice lat, double mass, double* x_dev, const double* xe_dev, const double* b_dev) {
  return guarded([&] {
    if (!x_dev || !xe_dev || !b_dev) throw ArgError("x_dev, xe_dev and b_dev must not be NULL");
    dev_eo_reconstruct(lattice_op(lat, OP_STAG, mass, 1), x_dev, xe_dev, b_dev, nullptr);
  });
}

int glx_set_tuning(int group, int use_graph) {
  return guarded([&] {
    if (group < 1 || group > 1024) throw ArgError("group must be in [1, 1024]");
    g_group = group;
    g_use_graph = use_graph ? 1 : 0;
  });
}

int glx_set_sstep(int s, int rows) {
  return guarded([&] { set_sstep(s, rows); });
}
int glx_get_sstep(int* s, int* rows) {
  return guarded([&] {
    if (s) *s = sstep_default();
    if (rows) *rows = sstep_rows();
  });
}
int glx_set_sstep_rows(int lx, int ly, int rows) {
  return guarded([&] {
    if (lx <= 0 || ly <= 0) throw ArgError("bad lattice size");
    if (rows > 0 && (rows < 8 || rows % 4 != 0)) throw ArgError("rows must be -1, 0 or a multiple of 4 >= 8");
    set_sstep_rows(lx, ly, rows < 0 ? -1 : rows);
  });
}
int glx_set_sstep_form(int form, int band_rows) {
  return guarded([&] { set_sstep_form(form, band_rows); });
}
int glx_get_sstep_form(int* form) {
  return guarded([&] {
    if (!form) throw ArgError("form is NULL");
    *form = sstep_form();
  });
}
int glx_set_sstep_form_entry(int lx, int ly, int form) {
  return guarded([&] {
    if (lx <= 0 || ly <= 0) throw ArgError("bad lattice size");
    set_sstep_form_entry(lx, ly, form);
  });
}
int glx_get_sstep_form_for(int lx, int ly, int* form) {
  return guarded([&] {
    if (!form) throw ArgError("form is NULL");
    if (lx <= 0 || ly <= 0) throw ArgError("bad lattice size");
    *form = sstep_form_for(lx, ly);
  });
}
int glx_get_sstep_rows(int lx, int ly, int* rows) {
  return guarded([&] {
    if (!rows) throw ArgError("rows is NULL");
    if (lx <= 0 || ly <= 0) throw ArgError("bad lattice size");
    *rows = sstep_rows_for(lx, ly);
  });
}

int glx_dev_cg_fused(glx_lattice lat, int kind, double mass, double* x_dev, const double* b_dev, int max_iter,
                     double eps, glx_verbose* verb, glx_inversion_info* info) {
  return guarded([&] {
    if (!x_dev || !b_dev) throw ArgError("x_dev and b_dev must not be NULL");
    DevOperator o = lattice_op(lat, kind, mass, 1);
    if (o.lx <= 0) throw ArgError("operator without links needs a lattice for its dimensions");
    Printer pr(verb);
    SolveOpts so;
    so.max_iter = max_iter;
    so.eps = eps;
    so.want_hist = pr.level() == GLX_VERB_DETAIL;
    so.group = g_group;
    so.use_graph = g_use_graph != 0;
    SolveResult r = dev_cg_fused(o, x_dev, b_dev, so);
    for (int k = 0; k < int(r.relres_hist.size()); k++) pr.resid("CG-fused", k + 1, 2 + k, r.relres_hist[k]);
    pr.summary("CG-fused", r.success, r.iter, r.ops_count, std::sqrt(r.resSq) / r.bsqrt);
    fill_info(info, r, "CG-fused");
  });
}

int glx_dev_apply(glx_lattice lat, int kind, double mass, int nc, int fp32, int link_prec, void* out_dev,
                  const void* in_dev) {
  return guarded([&] {
    if (!out_dev || !in_dev) throw ArgError("out_dev and in_dev must not be NULL");
    if (out_dev == in_dev) throw ArgError("out_dev and in_dev must be different buffers");
    DevOperator o = lattice_op(lat, kind, mass, nc);
    if (o.lx <= 0) throw ArgError("operator without links needs a lattice for its dimensions");
    dev_apply(o, out_dev, in_dev, link_prec, fp32, nullptr);
  });
}

// ------------------------------------------------------------ stream probe --
// Plain streaming kernels: what the device delivers to a kernel that only
// reads, only writes, copies, or reads three bytes for each byte it writes
// (the stencils' own ratio), with 16-byte accesses per lane in a grid-stride
// loop.  The ceiling the stencils' fractions of peak are read against.

int glx_dev_time_stream(int mode, size_t bytes, int reps, double* ms_per_pass, double* bytes_per_pass) {
  return guarded([&] {
    if (mode < 0 || mode > 3) throw ArgError("mode must be 0 (read), 1 (write), 2 (copy) or 3 (3 reads : 1 write)");
    if (reps <= 0 || !ms_per_pass) throw ArgError("bad arguments");
    const int arrays = mode == 3 ? 4 : (mode == 2 ? 2 : 1);
    const long n = long(bytes / (16 * size_t(arrays)));
    if (n < 1) throw ArgError("footprint too small");
    DevBuf buf(size_t(n) * 16 * arrays), sink(kBlock * 16);
    float4* v = buf.as<float4>();
    GLX_HIP(hipMemset(buf.get(), 0, size_t(n) * 16 * arrays));
    const float4* a = v;
    const float4* b = v + (arrays > 2 ? n : 0);
    const float4* c = v + (arrays > 2 ? 2 * n : 0);
    float4* o = v + (arrays - 1) * n;
    const int nb = int(std::min<long>((n + kBlock - 1) / kBlock, 8192));
    auto launch = [&]() {
      switch (mode) {
        case 0: hipLaunchKernelGGL(k_stream<0>, dim3(nb), dim3(kBlock), 0, 0, a, b, c, o, n, sink.as<float4>()); break;
        case 1: hipLaunchKernelGGL(k_stream<1>, dim3(nb), dim3(kBlock), 0, 0, a, b, c, o, n, sink.as<float4>()); break;
        case 2: hipLaunchKernelGGL(k_stream<2>, dim3(nb), dim3(kBlock), 0, 0, a, b, c, o, n, sink.as<float4>()); break;
        default: hipLaunchKernelGGL(k_stream<3>, dim3(nb), dim3(kBlock), 0, 0, a, b, c, o, n, sink.as<float4>());
      }
      GLX_CHECK_LAUNCH();
    };
    launch();  // warm-up
    hipEvent_t e0, e1;
    GLX_HIP(hipEventCreate(&e0));
    GLX_HIP(hipEventCreate(&e1));
    GLX_HIP(hipEventRecord(e0, 0));
    for (int k = 0; k < reps; k++) launch();
    GLX_HIP(hipEventRecord(e1, 0));
    GLX_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    GLX_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_per_pass = double(ms) / reps;
    if (bytes_per_pass) *bytes_per_pass = double(n) * 16.0 * arrays;
  });
}

int glx_dev_time_apply(glx_lattice lat, int kind, double mass, int fp32, int link_prec, void* out_dev,
                       const void* in_dev, int reps, double* ms_per_apply) {
  return guarded([&] {
    if (!out_dev || !in_dev || !ms_per_apply || reps <= 0) throw ArgError("bad arguments");
    DevOperator o = lattice_op(lat, kind, mass, 1);
    if (o.lx <= 0) throw ArgError("operator without links needs a lattice for its dimensions");
    dev_apply(o, out_dev, in_dev, link_prec, fp32, nullptr);  // warm-up
    hipEvent_t a, b;
    GLX_HIP(hipEventCreate(&a));
    GLX_HIP(hipEventCreate(&b));
    RedSpace red;
    red.init(kMaxPointBlocks, nullptr);
    GLX_HIP(hipDeviceSynchronize());
    GLX_HIP(hipEventRecord(a, nullptr));
    for (int k = 0; k < reps; k++) {
      if (fp32)
        run_stencil<c64>(nullptr, o, link_prec, InPlain<c64>{static_cast<const c64*>(in_dev)},
                         static_cast<c64*>(out_dev), EpiNone<c64>{}, FinNone{}, red.ws(), kNoGate);
      else if (kind == OP_LAPLACE_REAL)
        run_stencil<double>(nullptr, o, LINKS_F64, InPlain<double>{static_cast<const double*>(in_dev)},
                            static_cast<double*>(out_dev), EpiNone<double>{}, FinNone{}, red.ws(), kNoGate);
      else
        run_stencil<c128>(nullptr, o, LINKS_F64, InPlain<c128>{static_cast<const c128*>(in_dev)},
                          static_cast<c128*>(out_dev), EpiNone<c128>{}, FinNone{}, red.ws(), kNoGate);
    }
    GLX_HIP(hipEventRecord(b, nullptr));
    GLX_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    GLX_HIP(hipEventElapsedTime(&ms, a, b));
    *ms_per_apply = double(ms) / reps;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  });
}

int glx_set_ddwave_rows(int rows) {
  return guarded([&] {
    if (rows < -1 || rows > 4096) throw ArgError("rows must be -1 (by lattice size), 0 (off) or 1..4096");
    set_ddwave_rows(rows);
  });
}
int glx_get_ddwave_rows(int lx, int ly, int esize, int* rows) {
  return guarded([&] {
    if (!rows) throw ArgError("rows is NULL");
    if (lx <= 0 || ly <= 0) throw ArgError("bad lattice size");
    if (esize != 8 && esize != 16) throw ArgError("esize must be 8 (complex FP32) or 16 (complex FP64)");
    *rows = ddwave_fits(lx, ly) ? ddwave_rows_for(lx, ly, esize) : 0;
  });
}
int glx_set_dwave_rows(int rows) {
  return guarded([&] {
    if (rows < -1 || rows > 4096) throw ArgError("rows must be -1 (by lattice size), 0 (tiled) or 1..4096");
    set_dwave_rows(rows);
  });
}
int glx_get_dwave_rows(int lx, int ly, int esize, int* rows) {
  return guarded([&] {
    if (!rows) throw ArgError("rows is NULL");
    if (lx <= 0 || ly <= 0) throw ArgError("bad lattice size");
    if (esize != 8 && esize != 16) throw ArgError("esize must be 8 (complex FP32) or 16 (complex FP64)");
    *rows = dwave_fits(lx, ly) ? dwave_rows_for(lx, ly, esize) : 0;
  });
}
int glx_set_normal_rows(int lx, int ly, int esize, int rows) {
  return guarded([&] {
    if (lx < 0 || ly < 0 || (lx == 0) != (ly == 0)) throw ArgError("bad lattice size");
    if (rows < -1 || (rows > 0 && (rows < 8 || rows % 4 != 0)))
      throw ArgError("rows must be 0 or a multiple of 4 >= 8 (-1: the default by block count, or remove the "
                     "size's entry)");
    if (lx > 0 && esize != 8 && esize != 16) throw ArgError("esize must be 8 (complex FP32) or 16 (complex FP64)");
    set_normal_rows(lx, ly, esize, rows);
  });
}

int glx_get_normal_rows(int lx, int ly, int esize, int* rows) {
  return guarded([&] {
    if (!rows) throw ArgError("rows is NULL");
    *rows = normal_rows_for(lx, ly, esize);
  });
}

int glx_autotune_normal(glx_lattice lat, int reps, int* rows64, int* rows32, double* ms64, double* ms32) {
  return guarded([&] {
    if (!lat || reps <= 0) throw ArgError("bad arguments");
    // the timed work is the solvers' own inner loop: `reps` iterations of
    // the single-pass FP64 CG and of the mixed-precision CG (eps = 0, so
    // neither stops early) on b = 1, x0 = 0
    DevOperator o = lattice_op(lat, OP_STAG_NORMAL, 0.1, 1);
    const int n = o.size();
    std::vector<c128> ones(n, c128{1.0, 0.0});
    DevBuf b(size_t(n) * sizeof(c128)), x(size_t(n) * sizeof(c128));
    GLX_HIP(hipMemcpy(b.get(), ones.data(), size_t(n) * sizeof(c128), hipMemcpyHostToDevice));
    const int cands[] = {0, 8, 12, 16, 24, 32, 48, 64, 96, 128, 256};  // 256: ly >= 2048 only
    SolveOpts so;
    so.max_iter = reps;
    so.eps = 0.0;
    so.group = g_group;
    so.use_graph = g_use_graph != 0;
    MixedOpts mo;
    mo.link_prec = LINKS_F32;
    mo.sstep = 1;  // the table shapes the one-iteration kernel; the s-step kernel has its own rows
    auto time_one = [&](int fp32) {
      GLX_HIP(hipMemset(x.get(), 0, size_t(n) * sizeof(c128)));
      GLX_HIP(hipDeviceSynchronize());
      const auto t0 = std::chrono::steady_clock::now();
      if (fp32) (void)dev_cg_mixed(o, x.get(), b.get(), so, mo);
      else (void)dev_cg_fused(o, x.get(), b.get(), so);
      GLX_HIP(hipDeviceSynchronize());
      return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / reps;
    };
    for (int fp32 = 0; fp32 < 2; fp32++) {
      if (fp32 && !o.u32) continue;
      const int esize = fp32 ? 8 : 16;
      int best = 0;
      double best_ms = 1e300;
      const int saved = normal_rows_entry(o.lx, o.ly, esize);  // -1: none (erased again on error)
      try {
        for (int c : cands) {
          if (c == 256 && o.ly < 2048) continue;
          if (c > 0 && !normal_fits(o.lx, o.ly, c, fp32 ? normal_cpl<c64>() : normal_cpl<c128>())) continue;
          set_normal_rows(o.lx, o.ly, esize, c);
          time_one(fp32);  // warm-up
          const double t = std::min(time_one(fp32), time_one(fp32));
          if (t < best_ms) {
            best_ms = t;
            best = c;
          }
        }
      } catch (...) {
        set_normal_rows(o.lx, o.ly, esize, saved);
        throw;
      }
      set_normal_rows(o.lx, o.ly, esize, best);
      if (fp32) {
        if (rows32) *rows32 = best;
        if (ms32) *ms32 = best_ms;
      } else {
        if (rows64) *rows64 = best;
        if (ms64) *ms64 = best_ms;
      }
    }
  });
}

int glx_autotune_sstep(glx_lattice lat, int reps, int* rows, double* ms_per_iteration) {
  return guarded([&] {
    if (!lat || reps <= 0) throw ArgError("bad arguments");
    // `reps` iterations of the mixed-precision CG (eps = 0: no early stop) on
    // b = 1, x0 = 0, in its one-iteration form (candidate 0) and in s-step
    // form at each rows per block; the fastest per iteration is installed
    DevOperator o = lattice_op(lat, OP_STAG_NORMAL, 0.1, 1);
    if (!o.u32) throw ArgError("lattice was created without GLX_LINKS_F32");
    const int n = o.size();
    std::vector<c128> ones(n, c128{1.0, 0.0});
    DevBuf b(size_t(n) * sizeof(c128)), x(size_t(n) * sizeof(c128));
    GLX_HIP(hipMemcpy(b.get(), ones.data(), size_t(n) * sizeof(c128), hipMemcpyHostToDevice));
    SolveOpts so;
    so.max_iter = reps;
    so.eps = 0.0;
    so.group = g_group;
    so.use_graph = g_use_graph != 0;
    MixedOpts mo;
    mo.link_prec = LINKS_F32;
    mo.sstep = 2;
    const int forced = sstep_rows();
    const int saved = sstep_rows_entry(o.lx, o.ly);  // -1: none (erased again on error)
    const int forced_form = sstep_form(), saved_form = sstep_form_entry(o.lx, o.ly);
    set_sstep(sstep_default(), 0);  // the tables decide while timing
    set_sstep_form(-1, 0);
    int best = 0, best_form = 0;
    double best_ms = 1e300;
    auto restore = [&]() {
      set_sstep(sstep_default(), forced);
      set_sstep_form(forced_form, 0);
    };
    try {
      // candidates: (rows, form); the wave-strip form with the rows of the
      // size's default block shape (rows only route the solve to s-step form)
      std::vector<std::pair<int, int>> cand;
      for (int c : {0, 8, 16, 32, 64, 128, 256}) {
        if (c == 256 && o.ly < 2048) break;
        if (c > 0 && (o.lx % 128 != 0 || o.ly < 8 || o.ly % 4 != 0)) break;
        cand.push_back({c, 0});
      }
      if (cand.size() > 1) cand.push_back({sstep_rows_auto(o.lx, o.ly), 1});
      for (auto [c, f] : cand) {
        set_sstep_rows(o.lx, o.ly, c);
        set_sstep_form_entry(o.lx, o.ly, f);
        auto time_one = [&]() {
          GLX_HIP(hipMemset(x.get(), 0, size_t(n) * sizeof(c128)));
          GLX_HIP(hipDeviceSynchronize());
          const auto t0 = std::chrono::steady_clock::now();
          (void)dev_cg_mixed(o, x.get(), b.get(), so, mo);
          GLX_HIP(hipDeviceSynchronize());
          return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / reps;
        };
        time_one();  // warm-up
        const double t = std::min(time_one(), time_one());
        if (t < best_ms) {
          best_ms = t;
          best = c;
          best_form = f;
        }
      }
    } catch (...) {
      restore();
      set_sstep_rows(o.lx, o.ly, saved);
      set_sstep_form_entry(o.lx, o.ly, saved_form);
      throw;
    }
    restore();
    set_sstep_rows(o.lx, o.ly, best);
    set_sstep_form_entry(o.lx, o.ly, best_form);
    if (rows) *rows = best;
    if (ms_per_iteration) *ms_per_iteration = best_ms;
  });
}

int glx_dev_time_cg_kernel(glx_lattice lat, double mass, int prec, int reps, double* ms_per_launch,
                           double* bytes_per_site) {
  return guarded([&] {
    if (!ms_per_launch) throw ArgError("ms_per_launch is NULL");
    DevOperator o = lattice_op(lat, OP_STAG_NORMAL, mass, 1);
    *ms_per_launch = time_cg_kernel(o, prec, reps, bytes_per_site);
  });
}

int glx_dev_time_sscg_traffic(glx_lattice lat, int link_prec, int reps, double* ms_per_launch) {
  return guarded([&] {
    if (!ms_per_launch) throw ArgError("ms_per_launch is NULL");
    DevOperator o = lattice_op(lat, OP_STAG_NORMAL, 0.0, 1);
    *ms_per_launch = time_sscg_traffic_impl(o, link_prec == LINKS_F16 ? LINKS_F16 : LINKS_F32, reps);
  });
}

// ------------------------------------------------ general stencil handle --
struct glx_cstencil_s {
  glx_stencil_2d_nc desc{};  // dimensions, sdir, has_two, shifts (pointers cleared)
  DevOperator op;
  DevBuf coef, tmp;
};
namespace {
DevOperator cst_handle_op(glx_cstencil cs, int kind) {
  if (!cs) throw ArgError("stencil handle is NULL");
  if (!is_cst_kind(kind)) throw ArgError("kind must be one of GLX_OP_CSTENCIL*");
  DevOperator o = cs->op;
  o.kind = kind;
  return o;
}
}  // namespace

int glx_cstencil_create(glx_cstencil* out, const glx_stencil_2d_nc* st) {
  return guarded([&] {
    if (!out || !st) throw ArgError("cs and st must not be NULL");
    auto h = std::make_unique<glx_cstencil_s>();
    cst_upload(*st, h->op, h->coef, h->tmp);
    h->op.kind = OP_CST;
    h->desc = *st;
    h->desc.clover = h->desc.hopping = h->desc.two_link = nullptr;
    *out = h.release();
  });
}
int glx_cstencil_destroy(glx_cstencil cs) {
  return guarded([&] {
    GLX_HIP(hipDeviceSynchronize());
    delete cs;
  });
}
int glx_dev_cstencil_apply(glx_cstencil cs, int kind, int sdir, double* out_dev, const double* in_dev) {
  return guarded([&] {
    if (!out_dev || !in_dev) throw ArgError("out_dev and in_dev must not be NULL");
    if (sdir < 0 || sdir > 13) throw ArgError("sdir must be a stencil_dir (0..13)");
    DevOperator o = cst_handle_op(cs, kind);
    o.cs_sdir = sdir;
    dev_apply(o, out_dev, in_dev, LINKS_F64, 0, nullptr);
  });
}
int glx_dev_cstencil_solve(int solver, glx_cstencil cs, int kind, double* x_dev, const double* b_dev, int max_iter,
                           double eps, glx_verbose* verb, glx_inversion_info* info) {
  return guarded([&] {
    if (!x_dev || !b_dev) throw ArgError("x_dev and b_dev must not be NULL");
    DevOperator o = cst_handle_op(cs, kind);
    SolveResult r = solve_and_print(solver, o, x_dev, b_dev, max_iter, eps, verb);
    fill_info(info, r, solver_name(solver));
  });
}
int glx_dev_time_cstencil(glx_cstencil cs, int kind, double* out_dev, const double* in_dev, int reps,
                          double* ms_per_apply) {
  return guarded([&] {
    if (!out_dev || !in_dev || !ms_per_apply || reps <= 0) throw ArgError("bad arguments");
    DevOperator o = cst_handle_op(cs, kind);
    auto launch = [&] {
      launch_cst(nullptr, o, reinterpret_cast<c128*>(out_dev), reinterpret_cast<const c128*>(in_dev), kNoGate);
    };
    launch();
    GLX_HIP(hipDeviceSynchronize());
    hipEvent_t a, b;
    GLX_HIP(hipEventCreate(&a));
    GLX_HIP(hipEventCreate(&b));
    GLX_HIP(hipEventRecord(a, nullptr));
    for (int k = 0; k < reps; k++) launch();
    GLX_HIP(hipEventRecord(b, nullptr));
    GLX_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    GLX_HIP(hipEventElapsedTime(&ms, a, b));
    GLX_HIP(hipEventDestroy(a));
    GLX_HIP(hipEventDestroy(b));
    *ms_per_apply = double(ms) / reps;
  });
}

int glx_dev_alloc(void** ptr, size_t bytes) {
  return guarded([&] {
    if (!ptr) throw ArgError("ptr is NULL");
    GLX_HIP(hipMalloc(ptr, bytes));
  });
}
int glx_dev_free(void* ptr) {
  return guarded([&] { GLX_HIP(hipFree(ptr)); });
}
int glx_memcpy_htod(void* dst, const void* src, size_t bytes) {
  return guarded([&] { GLX_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); });
}
int glx_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
  return guarded([&] { GLX_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); });
}
int glx_memcpy_dtod(void* dst, const void* src, size_t bytes) {
  return guarded([&] { GLX_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice)); });
}
int glx_device_synchronize(void) {
  return guarded([&] { GLX_HIP(hipDeviceSynchronize()); });
}
int glx_device_count(int* count) {
  return guarded([&] {
    if (!count) throw ArgError("count is NULL");
    *count = 0;
    GLX_HIP(hipGetDeviceCount(count));
  });
}
int glx_release_cache(void) {
  return guarded([&] { DevicePool::get().trim(); });
}
const char* glx_last_error(void) { return g_last_error.c_str(); }
int glx_minv_unpreconditioned_r(double* lhs, double* rhs, int size, int type, const glx_inverter_params* params,
                                glx_matvec mv, void* extra, glx_verbose* verb, glx_inversion_info* info) {
  using P = int (*)(double*, double*, int, int, double, glx_matvec, void*, glx_verbose*, glx_inversion_info*);
  using R = int (*)(double*, double*, int, int, double, int, glx_matvec, void*, glx_verbose*, glx_inversion_info*);
  static const P plain[3] = {glx_minv_vector_cg_r, glx_minv_vector_bicgstab_r, glx_minv_vector_gcr_r};
  static const R restarted[3] = {glx_minv_vector_cg_restart_r, glx_minv_vector_bicgstab_restart_r,
                                 glx_minv_vector_gcr_restart_r};
  return dispatch_unpreconditioned(plain, restarted, lhs, rhs, size, type, params, mv, extra, verb, info);
}
int glx_minv_unpreconditioned_c(double* lhs, double* rhs, int size, int type, const glx_inverter_params* params,
                                glx_matvec mv, void* extra, glx_verbose* verb, glx_inversion_info* info) {
  using P = int (*)(double*, double*, int, int, double, glx_matvec, void*, glx_verbose*, glx_inversion_info*);
  using R = int (*)(double*, double*, int, int, double, int, glx_matvec, void*, glx_verbose*, glx_inversion_info*);
  static const P plain[3] = {glx_minv_vector_cg_c, glx_minv_vector_bicgstab_c, glx_minv_vector_gcr_c};
  static const R restarted[3] = {glx_minv_vector_cg_restart_c, glx_minv_vector_bicgstab_restart_c,
                                 glx_minv_vector_gcr_restart_c};
  return dispatch_unpreconditioned(plain, restarted, lhs, rhs, size, type, params, mv, extra, verb, info);
}

int glx_abi_version(void) { return GLX_ABI_VERSION; }

}  // extern "C"
