/*
 * ref_monte.c -- test-only driver around the compiled uFEM reference's Monte
 * plugin functions (fem/fem_mont.c built with USE_MONTE, as the reference's
 * libfem.so target does, fem/Makefile:192-193).
 *
 * ORACLE / TEST INFRASTRUCTURE.  Linked by oracle/Makefile (target
 * ref-monte) against objects compiled from the reference sources where they
 * lie; used only to write golden vectors of the plugin interface.  Nothing
 * in the product library links or calls it.
 *
 * Usage: ufem_ref_monte in.fem samples.txt
 *   samples.txt: one line of ilen input values per sample
 * Output (stdout): "ilen olen ffunc", then one line per input name and per
 * output name, then per sample "rv o_0 ... o_{olen-1}" (%.17g)
 */
#include <stdio.h>
#include <stdlib.h>

extern long monte_dlib_interface_type(void);
extern void monte_nums_of_vars(char *param, long *ilen, long *olen, long *ffunc);
extern char *monte_ivar_name(char *param, long pos);
extern char *monte_ovar_name(char *param, long pos);
extern int monte_init_lib_stuff(char *param);
extern int monte_clean_lib_stuff(char *param);
extern int monte_solution(char *param, double *ifld, double *ofld, long if_type);

int main(int argc, char **argv)
{
  long ilen = 0, olen = 0, ffunc = -1, i;
  double *ifld, *ofld;
  FILE *f;
  int rv;
  if (argc < 3) {
    fprintf(stderr, "usage: %s in.fem samples.txt\n", argv[0]);
    return 2;
  }
  if (monte_init_lib_stuff(argv[1]) != 0) {
    fprintf(stderr, "monte_init_lib_stuff failed\n");
    return 1;
  }
  monte_nums_of_vars(argv[1], &ilen, &olen, &ffunc);
  printf("%ld %ld %ld %ld\n", monte_dlib_interface_type(), ilen, olen, ffunc);
  for (i = 0; i < ilen; i++) printf("%s\n", monte_ivar_name(argv[1], i));
  for (i = 0; i < olen; i++) printf("%s\n", monte_ovar_name(argv[1], i));
  ifld = (double *)calloc((size_t)ilen + 1, sizeof(double));
  ofld = (double *)calloc((size_t)olen + 1, sizeof(double));
  if (!ifld || !ofld || !(f = fopen(argv[2], "r"))) return 1;
  for (;;) {
    for (i = 0; i < ilen; i++)
      if (fscanf(f, "%lf", &ifld[i]) != 1) break;
    if (i < ilen) break;
    rv = monte_solution(argv[1], ifld, ofld, 0);
    printf("%d", rv);
    for (i = 0; i < olen; i++) printf(" %.17g", ofld[i]);
    printf("\n");
    fflush(stdout);
  }
  fclose(f);
  monte_clean_lib_stuff(argv[1]);
  return 0;
}
