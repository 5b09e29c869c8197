// inverter_struct.h -- result record of every solver, as in the reference
// (inverter_struct.h:14-89): same members, same constructor, copy and
// assignment semantics (a copy duplicates resSqmrhs; assignment takes over
// the right-hand side's buffer).  Part of the drop-in C++ interface of the
// MI355X solver library; the solvers themselves run on the GPU (libglx.so).
#ifndef GLX_COMPAT_INVERTER_STRUCT
#define GLX_COMPAT_INVERTER_STRUCT

#include <iostream>
#include <string>
using namespace std;

struct inversion_info {
  double resSq;       // squared absolute true residual ||A x - b||^2
  int iter;           // iterations
  bool success;       // converged?
  std::string name;   // algorithm name, e.g. "CG" or "CG(8)"
  int ops_count;      // operator applications
  double* resSqmrhs;  // per-shift squared residuals (multishift solves only)
  int n_rhs;

  inversion_info() : resSq(0.0), iter(0), success(false), name(""), ops_count(0), resSqmrhs(0), n_rhs(-1) {}
  explicit inversion_info(int in_n_rhs)
      : resSq(0.0), iter(0), success(false), name(""), ops_count(0), resSqmrhs(new double[in_n_rhs]), n_rhs(in_n_rhs) {}
  ~inversion_info() {
    if (resSqmrhs && n_rhs > 0) delete[] resSqmrhs;
    resSqmrhs = 0;
  }
  inversion_info(const inversion_info& o)
      : resSq(o.resSq), iter(o.iter), success(o.success), name(o.name), ops_count(o.ops_count), resSqmrhs(0),
        n_rhs(o.n_rhs) {
    if (o.resSqmrhs && n_rhs > 0) {
      resSqmrhs = new double[n_rhs];
      for (int k = 0; k < n_rhs; k++) resSqmrhs[k] = o.resSqmrhs[k];
    }
  }
  inversion_info& operator=(inversion_info o) {  // by value: o is a private copy
    resSq = o.resSq;
    iter = o.iter;
    success = o.success;
    name = o.name;
    ops_count = o.ops_count;
    n_rhs = o.n_rhs;
    if (o.resSqmrhs && n_rhs > 0) {
      resSqmrhs = o.resSqmrhs;
      o.resSqmrhs = 0;
    }
    return *this;
  }
};

#endif
