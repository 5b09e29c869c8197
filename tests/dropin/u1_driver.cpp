// A driver written against the reference's u1_utils.h interface (the sequence
// of u1_utils/u1_test.cpp:17-83), built with g++ against the compat headers
// and linked with libglx.so.  Prints every result with 17 significant digits
// so the test can compare them bit for bit.
#include <cstdio>
#include <random>

#include "u1_utils.h"

static void show(const char* what, complex<double> p) { std::printf("%s %.17g %.17g\n", what, p.real(), p.imag()); }

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s <hot field file> <smeared field file>\n", argv[0]);
    return 2;
  }
  const int L = 16, V = L * L;
  std::mt19937 gen(1337u);
  complex<double>* f1 = new complex<double>[2 * V];
  complex<double>* f2 = new complex<double>[2 * V];
  complex<double>* g = new complex<double>[V];

  unit_gauge_u1(f1, L, L);
  show("unit", get_plaquette_u1(f1, L, L));
  for (double beta = 100; beta > 1e-4; beta *= 0.1) {
    gauss_gauge_u1(f1, L, L, gen, beta);
    char tag[64];
    std::snprintf(tag, sizeof tag, "gauss %.17g", beta);
    show(tag, get_plaquette_u1(f1, L, L));
    apply_ape_smear_u1(f2, f1, L, L, 0.0, 1);
    show("ape0", get_plaquette_u1(f2, L, L));
    std::printf("topo %.17g\n", get_topo_u1(f2, L, L));
  }
  rand_gauge_u1(f1, L, L, gen);
  show("hot", get_plaquette_u1(f1, L, L));
  write_gauge_u1(f1, L, L, argv[1]);
  read_gauge_u1(f2, L, L, argv[1]);
  show("loaded", get_plaquette_u1(f2, L, L));
  rand_trans_u1(g, L, L, gen);
  apply_gauge_trans_u1(f2, g, L, L);
  show("transformed", get_plaquette_u1(f2, L, L));

  // smearing: a cold-ish field, three rounds, before and after a transform
  gauss_gauge_u1(f1, L, L, gen, 4.0);
  apply_ape_smear_u1(f2, f1, L, L, 0.5, 3);
  show("ape3", get_plaquette_u1(f2, L, L));
  write_gauge_u1(f2, L, L, argv[2]);
  std::printf("ape3topo %.17g\n", get_topo_u1(f2, L, L));
  rand_trans_u1(g, L, L, gen);
  apply_gauge_trans_u1(f1, g, L, L);
  apply_ape_smear_u1(f2, f1, L, L, 0.5, 3);
  show("ape3g", get_plaquette_u1(f2, L, L));
  std::printf("ape3gtopo %.17g\n", get_topo_u1(f2, L, L));
  delete[] f1;
  delete[] f2;
  delete[] g;
  return 0;
}
