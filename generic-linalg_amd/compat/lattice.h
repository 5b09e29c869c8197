// lattice.h -- index arithmetic of the reference's Lattice class
// (lattice/lattice.h:4-191): same class, methods and results.
// Index of (coordinates x_0..x_{nd-1}, colour) is
//   ((x_{nd-1} L_{nd-2} + x_{nd-2}) ... ) L_0 + x_0) nc + colour,
// x_0 fastest.  index_is_even is true when sum(x) is even; coord_is_even, as
// in the reference, returns sum(x) % 2, i.e. true for ODD coordinates.
#ifndef GLX_COMPAT_LATTICE
#define GLX_COMPAT_LATTICE

#include <vector>

class Lattice {
 public:
  Lattice(int my_nd, int* my_lattice, int my_nc) : nd(my_nd), nc(my_nc), dims(my_lattice, my_lattice + my_nd) {
    volume = 1;
    for (int d : dims) volume *= d;
    lattice_size = volume * nc;
  }

  inline void coord_to_index(int& i, int* coord, int color) { i = coord_to_index(coord, color); }
  inline int coord_to_index(int* coord, int color) {
    int i = 0;
    for (int mu = nd - 1; mu >= 0; mu--) i = i * dims[mu] + coord[mu];
    return i * nc + color;
  }
  inline void index_to_coord(int i, int* coord, int& color) {
    color = i % nc;
    i /= nc;
    for (int mu = 0; mu < nd; mu++) {
      coord[mu] = i % dims[mu];
      i /= dims[mu];
    }
  }
  inline int index_to_coordinate_dir(int i, int mu) {
    i /= nc;
    for (int nu = 0; nu < mu; nu++) i /= dims[nu];
    return i % dims[mu];
  }
  inline int index_to_color(int i) { return i % nc; }
  inline bool index_is_even(int& i) {
    int s = 0, rest = i / nc;
    for (int mu = 0; mu < nd; mu++) {
      s += rest % dims[mu];
      rest /= dims[mu];
    }
    return (s + 1) % 2;
  }
  inline bool coord_is_even(int* coord) {
    int s = 0;
    for (int mu = 0; mu < nd; mu++) s += coord[mu];
    return s % 2;
  }
  inline void get_lattice(int* out) {
    for (int mu = 0; mu < nd; mu++) out[mu] = dims[mu];
  }
  inline int get_lattice_dimension(int mu) { return (mu >= 0 && mu < nd) ? dims[mu] : -1; }
  inline int get_nd() { return nd; }
  inline int get_nc() { return nc; }
  inline int get_volume() { return volume; }
  inline int get_lattice_size() { return lattice_size; }

 private:
  int nd;
  int nc;
  std::vector<int> dims;
  int volume;
  int lattice_size;
};

#endif
