// coarse_stencil.h -- the reference's stencil_2d (stencil_2d/coarse_stencil.h:
// 14-173): same enum, struct fields, constructors and functions.  The apply
// functions run on the GPU (libglx.so); passed to a solver as the operator,
// they are recognised and the whole solve runs on the GPU with the
// coefficients uploaded once.
//
// Coefficients (complex<double>, Nc x Nc blocks row-major per site):
//   clover[c + nc*i]                        i = lattice index (site, row)
//   hopping[c + nc*i + d*nc*lattice_size]   d: +x, +y, -x, -y
//   two_link[c + nc*i + d*nc*lattice_size]  d: +2x, +x+y, +2y, -x+y, -2x, -x-y, -2y, +x-y
#ifndef GLX_COMPAT_COARSE_STENCIL
#define GLX_COMPAT_COARSE_STENCIL

#include <algorithm>
#include <complex>
#include <iostream>

#include "generic_vector.h"
#include "lattice.h"

using namespace std;

enum stencil_dir {
  DIR_ALL = 0,
  DIR_0 = 1,
  DIR_XP1 = 2,
  DIR_YP1 = 3,
  DIR_XM1 = 4,
  DIR_YM1 = 5,
  DIR_XP2 = 6,
  DIR_XP1YP1 = 7,
  DIR_YP2 = 8,
  DIR_XM1YP1 = 9,
  DIR_XM2 = 10,
  DIR_XM1YM1 = 11,
  DIR_YM2 = 12,
  DIR_XP1YM1 = 13,
};

struct stencil_2d {
  Lattice* lat;
  stencil_dir sdir;
  complex<double>* clover;    // nc nc V
  complex<double>* hopping;   // nc nc V x {+x, +y, -x, -y}
  bool has_two;               // two-link terms present (stencil_size == 2)
  int stencil_size;
  complex<double>* two_link;  // nc nc V x 8, or null
  bool generated;
  complex<double> shift;      // identity shift
  complex<double> eo_shift;   // +shift on even sites, -shift on odd sites
  complex<double> dof_shift;  // +shift on colours < nc/2, -shift on the others

  stencil_2d(Lattice* in_lat, int in_stencil_size, complex<double> in_shift = 0.0,
             complex<double> in_eo_shift = 0.0, complex<double> in_dof_shift = 0.0)
      : lat(in_lat), sdir(DIR_ALL), has_two(in_stencil_size == 2), stencil_size(in_stencil_size), generated(false),
        shift(in_shift), eo_shift(in_eo_shift), dof_shift(in_dof_shift) {
    allocate();
  }
  stencil_2d(const stencil_2d& o)
      : lat(o.lat), sdir(o.sdir), has_two(o.has_two), stencil_size(o.stencil_size), generated(o.generated),
        shift(o.shift), eo_shift(o.eo_shift), dof_shift(o.dof_shift) {
    allocate();
    std::copy(o.clover, o.clover + block(), clover);
    std::copy(o.hopping, o.hopping + 4 * block(), hopping);
    if (has_two) std::copy(o.two_link, o.two_link + 8 * block(), two_link);
  }
  stencil_2d& operator=(const stencil_2d&) = delete;
  ~stencil_2d() {
    delete[] clover;
    delete[] hopping;
    delete[] two_link;
  }
  void clear_stencils() {
    std::fill(clover, clover + block(), complex<double>(0.0));
    std::fill(hopping, hopping + 4 * block(), complex<double>(0.0));
    if (has_two) std::fill(two_link, two_link + 8 * block(), complex<double>(0.0));
    generated = false;
  }

 private:
  int block() const { return lat->get_volume() * lat->get_nc() * lat->get_nc(); }
  void allocate() {
    clover = new complex<double>[block()]();
    hopping = new complex<double>[4 * block()]();
    two_link = has_two ? new complex<double>[8 * block()]() : nullptr;
  }
};

// lhs = stencil rhs (coarse_stencil.cpp:12-392); sdir selects one term
void apply_stencil_2d(complex<double>* lhs, complex<double>* rhs, void* extra_data);
// even output sites only, hopping terms (coarse_stencil.cpp:395-557)
void apply_stencil_2d_eo(complex<double>* lhs, complex<double>* rhs, void* extra_data);
// odd output sites only (coarse_stencil.cpp:560-722)
void apply_stencil_2d_oe(complex<double>* lhs, complex<double>* rhs, void* extra_data);
// colours < nc/2 from colours >= nc/2 (coarse_stencil.cpp:725-1117)
void apply_stencil_2d_tb(complex<double>* lhs, complex<double>* rhs, void* extra_data);
// colours >= nc/2 from colours < nc/2 (coarse_stencil.cpp:1120-1512)
void apply_stencil_2d_bt(complex<double>* lhs, complex<double>* rhs, void* extra_data);
// fill an allocated, zero stencil by applying matrix_vector to unit vectors
// (coarse_stencil.cpp:1515-1627); no-op if generated or stencil_size > 2
void generate_stencil_2d(stencil_2d* stenc, void (*matrix_vector)(complex<double>*, complex<double>*, void*),
                         void* extra_data);

#endif
