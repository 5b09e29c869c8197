/* The C ABI header is plain C: this file, compiled with gcc -std=c99
 * -pedantic -Werror, declares a value of every struct the header defines,
 * takes the address of a sample of entry points and calls the ones that
 * need no device. */
#include <stdio.h>

#include "glx.h"

int main(void) {
  glx_staggered_u1_op op = {0, 0.1, 8, 8, 1, 0.0};
  glx_verbose verb;
  glx_inversion_info info;
  glx_stencil_2d st1;
  glx_stencil_2d_nc st;
  glx_inverter_params params;
  glx_batch_system bs;
  glx_matvec mv = glx_square_staggered_normal_u1;
  int (*solve)(double*, double*, int, int, double, glx_matvec, void*, glx_verbose*, glx_inversion_info*) =
      glx_minv_vector_cg_c;
  int (*gen)(glx_stencil_2d_nc*, glx_matvec, void*) = glx_generate_stencil_2d_nc;
  (void)op;
  (void)verb;
  (void)info;
  (void)st1;
  (void)st;
  (void)params;
  (void)bs;
  printf("abi %d %d %d %d\n", glx_abi_version(), mv != 0, solve != 0, gen != 0);
  return 0;
}
