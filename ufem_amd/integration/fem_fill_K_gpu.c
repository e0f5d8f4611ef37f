/*
 * fem_fill_K_gpu.c -- the reference-side binding of the GPU stiffness fill.
 *
 * This file is built with uFEM's own sources (it includes fem/fem_sol.h and
 * reads the model through the globals of fem/fem_data.h) and defines
 * fem_fill_K, the function femSolve and the other solution drivers call to
 * assemble K (fem/fem_sol.c:897-989).  uFEM's own fem_fill_K stays in the
 * program under the name fem_fill_K_cpu: fem_sol.o is built with
 * -ffunction-sections, its fem_fill_K is made weak and an alias
 * fem_fill_K_cpu is added at the same address, so every call to fem_fill_K,
 * fem_sol.o's own included, comes here (oracle/Makefile, target fem-gpu).
 *
 * fem_fill_K(AF_NO) -- K += the element matrices -- runs on the GPU through
 * ufem_gpu_fill_K when every element is an 8-node brick (type 9,
 * fem/fem_e009.c) of material type 1 (linear isotropic, fem/fem_m001.c),
 * with the brick's 3 Gauss points per axis (fem/fem_e009.c:282), the
 * positions of K taken from nDOFfld as femLocKM_e takes them
 * (fem/fem_mloc.c:121-221).  Any other model, and fem_fill_K(AF_YES) (the
 * results pass: element forces, F, Fr and the reactions), goes to
 * fem_fill_K_cpu.  The environment variable UFEM_FILL_K=cpu sends every
 * call there.  A GPU error is returned to the caller as it is (no silent
 * CPU retry), with its message on msgout.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fem_sol.h"
#include "ufem_gpu.h"

extern long *nDOFfld;                 /* fem/fem_sol.c:48 */
extern int fem_fill_K_cpu(long mode); /* uFEM's fem_fill_K, fem/fem_sol.c:897 */

#define BRICK_TYPE 9
#define LINEAR_ISO 1

/* every element a brick of 8 nodes of material type 1 */
static int bricks_only(void)
{
  long e;
  if (eLen <= 0) return 0;
  for (e = 0; e < eLen; e++) {
    long mp = eMP[e];
    if (eType[e] != BRICK_TYPE || mp < 0 || mp >= mpLen || mpType[mp] != LINEAR_ISO) return 0;
    if (eFrom[e] < 0 || eFrom[e] + 8 > eLenL) return 0;
  }
  return 1;
}

static int gpu_fill(void)
{
  long e, k, *enodes;
  double *E, *nu, t[3];
  ufem_mesh m;
  int rv;

  enodes = (long *)malloc(sizeof(long) * 8 * (size_t)eLen);
  E = (double *)malloc(sizeof(double) * (size_t)eLen);
  nu = (double *)malloc(sizeof(double) * (size_t)eLen);
  if (!enodes || !E || !nu) {
    free(enodes);
    free(E);
    free(nu);
    return AF_ERR_MEM;
  }
  for (e = 0; e < eLen; e++) {
    for (k = 0; k < 8; k++) enodes[8 * e + k] = eNodesL[eFrom[e] + k];
    E[e] = femGetMPValPos(e, MAT_EX, 0);
    nu[e] = femGetMPValPos(e, MAT_NU, 0);
  }
  memset(&m, 0, sizeof(m));
  m.n_nodes = nLen;
  m.x = n_x;
  m.y = n_y;
  m.z = n_z;
  m.n_elems = eLen;
  m.enodes = enodes;
  m.E = E;
  m.nu = nu;
  rv = ufem_gpu_fill_K(&K, &m, nDOFfld, KNOWN_DOFS, 3);
  free(enodes);
  free(E);
  free(nu);
  if (rv != AF_OK) {
    fprintf(msgout, "[E] %s: %s\n", "stiffness matrix on the GPU", ufem_gpu_last_error());
    return rv;
  }
  ufem_gpu_fill_K_timings(t);
#ifdef RUN_VERBOSE
  fprintf(msgout, "[i]     stiffness matrix on the GPU: %li bricks, set-up %.1f ms, device %.1f ms, "
          "write-back %.1f ms\n", eLen, t[0], t[1], t[2]);
#endif
  return AF_OK;
}

int fem_fill_K(long mode)
{
  const char *env = getenv("UFEM_FILL_K");
  if (mode != AF_NO || (env && strcmp(env, "cpu") == 0) || !bricks_only()) return fem_fill_K_cpu(mode);
  return gpu_fill();
}
