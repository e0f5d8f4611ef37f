/*
 * ref_harness.c -- test-only driver around the compiled uFEM reference.
 *
 * ORACLE / TEST INFRASTRUCTURE.  This file is linked against objects compiled
 * from the reference sources (uFEM 0.5.3, fem/*.c) by oracle/Makefile into
 * oracle/_ref/.  It is only ever used to write golden vectors and to validate
 * the clean-room CPU restatement (oracle/ufem_oracle.c); nothing in the
 * product library links or calls it.
 *
 * It replays the linear-static sequence of femSolve (fem/fem_sol.c:1655-1817)
 * step by step so that intermediate state can be dumped at full precision
 * (the reference's own writers print %e, fem/fem_io.c:689,697):
 *
 *   femReadInput          fem/fem_io.c:380      (preparsed input, -p)
 *   femElemTypeInit       fem/fem_eini.c:56
 *   femMatTypeInit        fem/fem_mini.c:43
 *   fem_dofs              fem/fem_sol.c:1263
 *   fem_sol_alloc         fem/fem_sol.c:485
 *   fem_sol_res_alloc     fem/fem_sol.c:639
 *   fem_fill_K(AF_NO)     fem/fem_sol.c:897     -> dump K0 (row-slot, slot order)
 *   fem_add_loads(0)      fem/fem_sol.c:1119
 *   fem_add_disps(AF_YES) fem/fem_sol.c:1195    -> dump K1, F
 *   femEqsCGwJ / femEqsBiCCSwJ  fem/fem_eqs.c:88 / :262
 *                                               -> dump u, SpMV call count
 *   fem_fill_K(AF_YES)    fem/fem_sol.c:897     -> dump reactions resRval
 *   femWriteRes           fem/fem_io.c:648      -> outdir/result.res
 *
 * SpMV calls made by the solvers are counted through the linker's
 * --wrap=femMatVecMultBig (fem_eqs.o references the symbol defined in
 * fem_math.o), which gives the iteration count the reference never reports.
 *
 * Usage:
 *   ufem_ref_harness -i in.fem -o outdir [-solver cg|bicgs] [-eps E]
 *                    [-maxit N] [-t THREADS] [-ke ePos ...] [-nodump-k]
 *                    [-fbc] [-np]
 *   -fbc  zero-valued supports left out of K (femFastBC, the fem binary's
 *         -fbc, fem/fem_para.c:1053-1057)
 *   -np   the input carries IDs, renumbered to positions by femRenumIndex
 *         (fem/fem_io.c:50-80; the fem binary without -p); the renumbered
 *         element nodes, load nodes and material/real-set indices are dumped
 * Output: raw little-endian arrays outdir/<name>.i64 / .f64 and a JSON
 * summary on stdout (phase wall times, iteration count, return codes).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <sys/stat.h>

#include "fem.h"
#include "fem_math.h"
#include "fem_data.h"
#include "fem_sol.h"
#include "fem_elem.h"
#include "fem_mat.h"
#include "fem_dama.h"

/* entry points of fem_sol.c that have no prototype in its header */
extern void fem_sol_null(void);
extern int fem_dofs(void);
extern int fem_sol_alloc(void);
extern int fem_sol_res_alloc(void);
extern int fem_fill_K(long mode);
extern int fem_add_loads(long step);
extern int fem_add_disps(long disp_mode, long step);
extern long femPreparsedData;
extern long femFastBC;
extern long femUseThreads;
extern long femThreadNum;
extern long femThreadMin;
extern long nDOFAct;
extern long nDOFlen;
extern long *nDOFfld;
extern tElem *Elem;

/* --wrap counter around the row-slot SpMV (fem/fem_math.c:1350) */
extern int __real_femMatVecMultBig(tMatrix *a, tVector *b, tVector *c);
static long g_spmv_calls = 0;
int __wrap_femMatVecMultBig(tMatrix *a, tVector *b, tVector *c)
{
  g_spmv_calls++;
  return __real_femMatVecMultBig(a, b, c);
}

static double now_s(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static const char *g_out = ".";

static void dump_raw(const char *name, const char *ext, const void *p, size_t esz, size_t n)
{
  char path[4096];
  FILE *f;
  snprintf(path, sizeof(path), "%s/%s.%s", g_out, name, ext);
  f = fopen(path, "wb");
  if (f == NULL) { fprintf(stderr, "cannot write %s\n", path); exit(3); }
  if (n > 0 && fwrite(p, esz, n, f) != n) { fprintf(stderr, "short write %s\n", path); exit(3); }
  fclose(f);
}

/* Row-slot matrix (fem/fem_math.h:47-57): dump per-row filled-slot counts,
 * the 1-based column ids in slot (insertion) order and the values. */
static void dump_rowslot(const char *tag, tMatrix *K)
{
  long i, j, nfill = 0, k = 0;
  long *rowlen, *cols;
  double *vals;
  char name[256];
  rowlen = (long *)malloc(sizeof(long) * (size_t)K->rows);
  for (i = 0; i < K->rows; i++) {
    long c = 0;
    for (j = K->frompos[i]; j < K->frompos[i] + K->defpos[i]; j++) {
      if (K->pos[j] <= 0) break;
      c++;
    }
    rowlen[i] = c;
    nfill += c;
  }
  cols = (long *)malloc(sizeof(long) * (size_t)(nfill + 1));
  vals = (double *)malloc(sizeof(double) * (size_t)(nfill + 1));
  for (i = 0; i < K->rows; i++) {
    for (j = K->frompos[i]; j < K->frompos[i] + rowlen[i]; j++) {
      cols[k] = K->pos[j];
      vals[k] = K->data[j];
      k++;
    }
  }
  snprintf(name, sizeof(name), "%s_rowlen", tag); dump_raw(name, "i64", rowlen, sizeof(long), (size_t)K->rows);
  snprintf(name, sizeof(name), "%s_defpos", tag); dump_raw(name, "i64", K->defpos, sizeof(long), (size_t)K->rows);
  snprintf(name, sizeof(name), "%s_cols", tag); dump_raw(name, "i64", cols, sizeof(long), (size_t)nfill);
  snprintf(name, sizeof(name), "%s_vals", tag); dump_raw(name, "f64", vals, sizeof(double), (size_t)nfill);
  free(rowlen); free(cols); free(vals);
}

int main(int argc, char **argv)
{
  const char *fin = NULL, *solver = "cg";
  double eps = FEM_ZERO / 10000.0;    /* value femSolve passes, fem/fem_sol.c:1722 */
  long maxit = -1, threads = 0, i;
  long ke_list[64]; int n_ke = 0;
  int dumpk = 1, rv, rv_solve = 0, fbc = 0, preparsed = 1;
  double t0, t_read, t_dofs, t_fill, t_bc, t_solve, t_react;
  double normA = 0.0;

  for (i = 1; i < argc; i++) {
    if (!strcmp(argv[i], "-i") && i + 1 < argc) fin = argv[++i];
    else if (!strcmp(argv[i], "-o") && i + 1 < argc) g_out = argv[++i];
    else if (!strcmp(argv[i], "-solver") && i + 1 < argc) solver = argv[++i];
    else if (!strcmp(argv[i], "-eps") && i + 1 < argc) eps = atof(argv[++i]);
    else if (!strcmp(argv[i], "-maxit") && i + 1 < argc) maxit = atol(argv[++i]);
    else if (!strcmp(argv[i], "-t") && i + 1 < argc) threads = atol(argv[++i]);
    else if (!strcmp(argv[i], "-ke") && i + 1 < argc && n_ke < 64) ke_list[n_ke++] = atol(argv[++i]);
    else if (!strcmp(argv[i], "-nodump-k")) dumpk = 0;
    else if (!strcmp(argv[i], "-fbc")) fbc = 1;
    else if (!strcmp(argv[i], "-np")) preparsed = 0;
    else { fprintf(stderr, "bad arg %s\n", argv[i]); return 2; }
  }
  if (fin == NULL) { fprintf(stderr, "need -i\n"); return 2; }
  mkdir(g_out, 0755);

  msgout = stderr;
  femPreparsedData = preparsed ? AF_YES : AF_NO;   /* -p, fem/fem_para.c:275 */
  femFastBC = fbc ? AF_YES : AF_NO;                 /* -fbc, fem/fem_para.c:1053 */
  if (threads > 0) {                  /* -t N, fem/fem_para.c:584-606 */
    femUseThreads = AF_YES;
    femThreadNum = threads > AF_MAX_THREADS ? AF_MAX_THREADS : threads;
  } else {
    femUseThreads = AF_NO;
  }

  t0 = now_s();
  if ((rv = femDataNull()) != AF_OK) return 10;
  if ((rv = femResNull()) != AF_OK) return 11;
  if ((rv = femReadInput((char *)fin)) != AF_OK) { fprintf(stderr, "read %d\n", rv); return 12; }
  t_read = now_s();
  if (!preparsed) {
    dump_raw("eNodesL", "i64", eNodesL, sizeof(long), (size_t)eLenL);
    dump_raw("eMP", "i64", eMP, sizeof(long), (size_t)eLen);
    dump_raw("eRS", "i64", eRS, sizeof(long), (size_t)eLen);
    if (nlLen > 0) dump_raw("nlNode", "i64", nlNode, sizeof(long), (size_t)nlLen);
  }

  if ((rv = femElemTypeInit()) != AF_OK) return 13;
  if ((rv = femMatTypeInit()) != AF_OK) return 14;
  fem_sol_null();
  femResNull();
  if ((rv = fem_dofs()) != AF_OK) { fprintf(stderr, "dofs %d\n", rv); return 15; }
  if ((rv = fem_sol_alloc()) != AF_OK) return 16;
  if ((rv = fem_sol_res_alloc()) != AF_OK) return 17;
  t_dofs = now_s();

  dump_raw("nDOFfld", "i64", nDOFfld, sizeof(long), (size_t)nDOFlen);

  /* element stiffness matrices on request (Mode = AF_NO) */
  for (i = 0; i < n_ke; i++) {
    tMatrix Ke; char name[64];
    femMatNull(&Ke);
    femFullMatInit(&Ke, 24, 24);
    rv = Elem[9].stiff(ke_list[i], AF_NO, &Ke, NULL, NULL);
    if (rv != AF_OK) { fprintf(stderr, "stiff %d\n", rv); return 18; }
    snprintf(name, sizeof(name), "Ke_%ld", ke_list[i]);
    dump_raw(name, "f64", Ke.data, sizeof(double), 576);
    femMatFree(&Ke);
  }

  if ((rv = fem_fill_K(AF_NO)) != AF_OK) { fprintf(stderr, "fill %d\n", rv); return 19; }
  t_fill = now_s();
  if (dumpk) dump_rowslot("K0", &K);

  if ((rv = fem_add_loads(0)) != AF_OK) return 20;
  dump_raw("F0", "f64", F.data, sizeof(double), (size_t)F.len);
  if ((rv = fem_add_disps(AF_YES, 0)) != AF_OK) return 21;
  t_bc = now_s();
  if (dumpk) dump_rowslot("K1", &K);
  dump_raw("F1", "f64", F.data, sizeof(double), (size_t)F.len);
  normA = femMatNormBig(&K);

  if (maxit < 0) maxit = nDOFAct;     /* femSolve passes nDOFAct, fem/fem_sol.c:1722 */
  g_spmv_calls = 0;
  if (!strcmp(solver, "bicgs")) rv_solve = femEqsBiCCSwJ(&K, &F, &u, eps, maxit);
  else rv_solve = femEqsCGwJ(&K, &F, &u, eps, maxit);
  t_solve = now_s();
  dump_raw("u", "f64", u.data, sizeof(double), (size_t)u.len);

  if ((rv = fem_fill_K(AF_YES)) != AF_OK) { fprintf(stderr, "fillY %d\n", rv); return 22; }
  t_react = now_s();
  if (resRLen > 0) {
    dump_raw("R", "f64", resRval, sizeof(double), (size_t)resRLen);
    dump_raw("Rnode", "i64", resRnode, sizeof(long), (size_t)resRLen);
    dump_raw("Rdof", "i64", resRdof, sizeof(long), (size_t)resRLen);
  }
  {
    /* the reference's own result file (fem/fem_io.c:648-734) */
    char path[4096];
    snprintf(path, sizeof(path), "%s/result.res", g_out);
    if ((rv = femWriteRes(path)) != AF_OK) { fprintf(stderr, "writeres %d\n", rv); return 23; }
  }

  printf("{\"n\": %ld, \"solver\": \"%s\", \"rv\": %d, \"spmv_calls\": %ld, "
         "\"normA\": %.17g, \"eps\": %.17g, \"maxit\": %ld, \"threads\": %ld, "
         "\"t_read\": %.6f, \"t_dofs\": %.6f, \"t_fill\": %.6f, \"t_bc\": %.6f, "
         "\"t_solve\": %.6f, \"t_react\": %.6f, \"fbc\": %d, \"preparsed\": %d}\n",
         nDOFAct, solver, rv_solve, g_spmv_calls, normA, eps, maxit, threads,
         t_read - t0, t_dofs - t_read, t_fill - t_dofs, t_bc - t_fill,
         t_solve - t_bc, t_react - t_solve, fbc, preparsed);
  return 0;
}
