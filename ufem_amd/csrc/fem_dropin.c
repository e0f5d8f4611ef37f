/*
 * fem_dropin.c -- the drop-in entry points with uFEM's own signatures
 * (include/ufem_fem_math.h): the two Jacobi solvers of fem/fem_eqs.c and the
 * "Big" sparse kernels of fem/fem_math.c, served by the GPU.
 *
 * A row-slot matrix (fem/fem_math.h:47-57) is read as the reference reads it:
 * row i owns slots frompos[i] .. frompos[i]+defpos[i]-1 and ends at the first
 * slot whose 1-based column is <= 0 (fem/fem_math.c:1424-1433).  It is copied
 * to the device in SELL-64 form with scalar blocks (columns sorted per row) on
 * every call, since uFEM mutates K in host memory between calls.
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ufem_fem_math.h"
#include "ufem_internal.h"

static int g_rule = 0;
static ufem_solve_info g_last;

void ufem_gpu_set_stop_rule(int rule) { g_rule = rule ? 1 : 0; }
void ufem_gpu_last_solve(ufem_solve_info *info)
{
  if (info) *info = g_last;
}

static int gpu_fail(const char *what)
{
  ufem_set_error("%s: %s", what, ugk_error());
  return AF_ERR;
}

/* one device context for the drop-in calls, grown on demand */
typedef struct {
  int ready;
  void *stream, *ev0, *ev1;
  ugk_mat K;
  long cap_slots, cap_rows;
  long n;             /* vectors sized for n */
  double *d_b, *d_x, *d_y;
  ufem_work w;
} dropin_ctx;

static dropin_ctx g_ctx;

static int ctx_ready(void)
{
  if (g_ctx.ready) return AF_OK;
  if (ugk_init() != 0) return gpu_fail("no usable GPU");
  g_ctx.stream = ugk_stream_create();
  g_ctx.ev0 = ugk_event_create();
  g_ctx.ev1 = ugk_event_create();
  if (!g_ctx.stream || !g_ctx.ev0 || !g_ctx.ev1) return gpu_fail("stream");
  g_ctx.ready = 1;
  return AF_OK;
}

static int ctx_vectors(long n, int solver)
{
  if (g_ctx.n >= n && g_ctx.w.n >= n && (solver != 1 || g_ctx.w.rhat)) return AF_OK;
  ugk_free(g_ctx.d_b); ugk_free(g_ctx.d_x); ugk_free(g_ctx.d_y);
  ufem_work_free(&g_ctx.w);
  g_ctx.d_b = (double *)ugk_malloc(sizeof(double) * (size_t)n);
  g_ctx.d_x = (double *)ugk_malloc(sizeof(double) * (size_t)n);
  g_ctx.d_y = (double *)ugk_malloc(sizeof(double) * (size_t)n);
  if (!g_ctx.d_b || !g_ctx.d_x || !g_ctx.d_y) { g_ctx.n = 0; return gpu_fail("vectors"); }
  g_ctx.n = n;
  if (ufem_work_alloc(&g_ctx.w, n, 1, 0) != AF_OK) { g_ctx.n = 0; return AF_ERR_MEM; }
  return AF_OK;
}

typedef struct { long c; double v; } colval;
static int cmp_cv(const void *a, const void *b)
{
  long x = ((const colval *)a)->c, y = ((const colval *)b)->c;
  return (x > y) - (x < y);
}

/* row-slot -> device SELL-64 (scalar blocks) */
static int upload_matrix(const tMatrix *a)
{
  long n = a->rows, nsl = (n + 63) / 64, i, total = 0, nnz = 0;
  int *rowlen = NULL, *slen = NULL, *cols = NULL;
  long *soff = NULL;
  double *vals = NULL;
  colval *buf = NULL;
  long maxrow = 0;
  int rv = AF_ERR_MEM;
  rowlen = (int *)calloc((size_t)n, sizeof(int));
  slen = (int *)calloc((size_t)nsl, sizeof(int));
  soff = (long *)calloc((size_t)nsl, sizeof(long));
  if (!rowlen || !slen || !soff) goto out;
  for (i = 0; i < n; i++) {
    long j, c = 0;
    for (j = a->frompos[i]; j < a->frompos[i] + a->defpos[i]; j++) {
      if (a->pos[j] <= 0) break;
      if (a->pos[j] > a->cols) { ufem_set_error("column %ld out of range in row %ld", a->pos[j], i + 1); rv = AF_ERR_SIZ; goto out; }
      c++;
    }
    rowlen[i] = (int)c;
    nnz += c;
    if (c > maxrow) maxrow = c;
  }
  for (i = 0; i < nsl; i++) {
    int l, mx = 0;
    for (l = 0; l < 64; l++) if (i * 64 + l < n && rowlen[i * 64 + l] > mx) mx = rowlen[i * 64 + l];
    slen[i] = mx;
    soff[i] = total;
    total += mx;
  }
  cols = (int *)malloc(sizeof(int) * (size_t)(total * 64 + 1));
  vals = (double *)malloc(sizeof(double) * (size_t)(total * 64 + 1));
  buf = (colval *)malloc(sizeof(colval) * (size_t)(maxrow + 1));
  if (!cols || !vals || !buf) goto out;
  for (i = 0; i < nsl; i++) {
    int l;
    for (l = 0; l < 64; l++) {
      long I = i * 64 + l, k;
      long L = I < n ? rowlen[I] : 0;
      if (I < n) {
        for (k = 0; k < L; k++) { buf[k].c = a->pos[a->frompos[I] + k] - 1; buf[k].v = a->data[a->frompos[I] + k]; }
        qsort(buf, (size_t)L, sizeof(colval), cmp_cv);
        for (k = 1; k < L; k++)
          if (buf[k].c == buf[k - 1].c) { ufem_set_error("duplicate column in row %ld", I + 1); rv = AF_ERR_VAL; goto out; }
      }
      for (k = 0; k < slen[i]; k++) {
        long at = (soff[i] + k) * 64 + l;
        if (k < L) { cols[at] = (int)buf[k].c; vals[at] = buf[k].v; }
        else { cols[at] = I < n ? (int)I : 0; vals[at] = 0.0; }
      }
    }
  }
  if (total > g_ctx.cap_slots || nsl > g_ctx.cap_rows / 64 + 1 || n > g_ctx.cap_rows) {
    ugk_free(g_ctx.K.d_slen); ugk_free(g_ctx.K.d_soff); ugk_free(g_ctx.K.d_rowlen);
    ugk_free(g_ctx.K.d_cols); ugk_free(g_ctx.K.d_vals);
    memset(&g_ctx.K, 0, sizeof(g_ctx.K));
    g_ctx.K.d_slen = (int *)ugk_malloc(sizeof(int) * (size_t)(nsl + 1));
    g_ctx.K.d_soff = (long *)ugk_malloc(sizeof(long) * (size_t)(nsl + 1));
    g_ctx.K.d_rowlen = (int *)ugk_malloc(sizeof(int) * (size_t)(n + 1));
    g_ctx.K.d_cols = (int *)ugk_malloc(sizeof(int) * (size_t)(total * 64 + 64));
    g_ctx.K.d_vals = (double *)ugk_malloc(sizeof(double) * (size_t)(total * 64 + 64));
    if (!g_ctx.K.d_slen || !g_ctx.K.d_soff || !g_ctx.K.d_rowlen || !g_ctx.K.d_cols || !g_ctx.K.d_vals) {
      g_ctx.cap_slots = g_ctx.cap_rows = 0;
      rv = gpu_fail("matrix allocation");
      goto out;
    }
    g_ctx.cap_slots = total;
    g_ctx.cap_rows = n;
  }
  g_ctx.K.bs = 1;
  g_ctx.K.gen++;
  g_ctx.K.n_brows = n;
  g_ctx.K.n_slices = nsl;
  g_ctx.K.total_slots = total;
  g_ctx.K.nnzb = nnz;
  if (ugk_h2d(g_ctx.K.d_slen, slen, sizeof(int) * (size_t)nsl, g_ctx.stream) ||
      ugk_h2d(g_ctx.K.d_soff, soff, sizeof(long) * (size_t)nsl, g_ctx.stream) ||
      ugk_h2d(g_ctx.K.d_rowlen, rowlen, sizeof(int) * (size_t)n, g_ctx.stream) ||
      ugk_h2d(g_ctx.K.d_cols, cols, sizeof(int) * (size_t)total * 64, g_ctx.stream) ||
      ugk_h2d(g_ctx.K.d_vals, vals, sizeof(double) * (size_t)total * 64, g_ctx.stream) ||
      ugk_sync(g_ctx.stream)) {
    rv = gpu_fail("matrix upload");
    goto out;
  }
  rv = AF_OK;
out:
  free(rowlen); free(slen); free(soff); free(cols); free(vals); free(buf);
  return rv;
}

static int check_mv(tMatrix *a, tVector *b, tVector *x)
{
  if (a->type != MAT_SPAR || b->type != VEC_FULL || x->type != VEC_FULL) {
    ufem_set_error("GPU path serves MAT_SPAR matrices and VEC_FULL vectors");
    return AF_ERR_TYP;
  }
  if (a->cols != x->rows || x->rows != b->rows || a->rows != b->rows) return AF_ERR_SIZ;
  return AF_OK;
}

static int dropin_solve(tMatrix *a, tVector *b, tVector *x, double eps, long maxIt, int solver)
{
  ufem_solve_opts o;
  int rv;
  long n = a->rows;
  memset(&g_last, 0, sizeof(g_last));
  if ((rv = check_mv(a, b, x)) != AF_OK) return rv;
  if (n == 0) return AF_OK;
  if ((rv = ctx_ready()) != AF_OK) return rv;
  if ((rv = upload_matrix(a)) != AF_OK) return rv;
  if ((rv = ctx_vectors(n, solver)) != AF_OK) return rv;
  if (ugk_h2d(g_ctx.d_b, b->data, sizeof(double) * (size_t)n, g_ctx.stream) ||
      ugk_h2d(g_ctx.d_x, x->data, sizeof(double) * (size_t)n, g_ctx.stream))
    return gpu_fail("vector upload");
  memset(&o, 0, sizeof(o));
  o.solver = solver;
  o.rule = g_rule;
  o.eps = eps;
  o.maxit = maxIt > 0 ? maxIt : -1;   /* maxIt <= 0: the loop runs zero times */
  o.exec = UFEM_EXEC_STREAMED;
  rv = ufem_solve_device(&g_ctx.K, n, g_ctx.d_b, g_ctx.d_x, &g_ctx.w, g_ctx.stream, g_ctx.ev0, g_ctx.ev1,
                         &o, &g_last);
  if (rv == AF_ERR_ZER || rv == AF_ERR_MEM) return rv;
  if (ugk_d2h(x->data, g_ctx.d_x, sizeof(double) * (size_t)n, g_ctx.stream)) return gpu_fail("x download");
  return rv;
}

/* femEqsCGwJ, fem/fem_eqs.c:88-247 */
int femEqsCGwJ(tMatrix *a, tVector *b, tVector *x, double eps, long maxIt)
{
  return dropin_solve(a, b, x, eps, maxIt, UFEM_SOLVER_PCG);
}

/* femEqsCGwSSOR, fem/fem_eqs.c:999-1188 (colour-ordered sweeps, see header) */
int femEqsCGwSSOR(tMatrix *a, tVector *b, tVector *x, double eps, long maxIt)
{
  return dropin_solve(a, b, x, eps, maxIt, UFEM_SOLVER_SSOR);
}

/* femEqsBiCCSwJ, fem/fem_eqs.c:262-468 (true-residual test, see header) */
int femEqsBiCCSwJ(tMatrix *a, tVector *b, tVector *x, double eps, long maxIt)
{
  return dropin_solve(a, b, x, eps, maxIt, UFEM_SOLVER_BICGSTAB);
}

/* femMatVecMultBig, fem/fem_math.c:1350-1451 */
int femMatVecMultBig(tMatrix *a, tVector *b, tVector *c)
{
  int rv;
  long n;
  if (a->cols != b->rows || c->rows != a->rows) return AF_ERR_SIZ;
  if (c->type != VEC_FULL) return AF_ERR_VAL;
  if (a->type != MAT_SPAR || b->type != VEC_FULL) {
    ufem_set_error("GPU path serves MAT_SPAR x VEC_FULL");
    return AF_ERR_TYP;
  }
  n = a->rows;
  if (n == 0) return AF_OK;
  if ((rv = ctx_ready()) != AF_OK) return rv;
  if ((rv = upload_matrix(a)) != AF_OK) return rv;
  if ((rv = ctx_vectors(n > b->rows ? n : b->rows, 0)) != AF_OK) return rv;
  if (ugk_h2d(g_ctx.d_x, b->data, sizeof(double) * (size_t)b->rows, g_ctx.stream) ||
      ugk_spmv(&g_ctx.K, g_ctx.d_x, g_ctx.d_y, g_ctx.stream) ||
      ugk_d2h(c->data, g_ctx.d_y, sizeof(double) * (size_t)n, g_ctx.stream))
    return gpu_fail("spmv");
  return AF_OK;
}

static int vec_upload2(tVector *a, tVector *b)
{
  int rv;
  long n = a->rows;
  if ((rv = ctx_ready()) != AF_OK) return rv;
  if ((rv = ctx_vectors(n, 0)) != AF_OK) return rv;
  if (ugk_h2d(g_ctx.d_x, a->data, sizeof(double) * (size_t)n, g_ctx.stream)) return gpu_fail("upload");
  if (b && ugk_h2d(g_ctx.d_y, b->data, sizeof(double) * (size_t)n, g_ctx.stream)) return gpu_fail("upload");
  return AF_OK;
}

static double dot_dev(long n, const double *x, const double *y)
{
  double r = 0.0;
  if (ugk_dot(n, x, y, g_ctx.w.q, g_ctx.w.part, g_ctx.stream) ||
      ugk_d2h(&r, g_ctx.w.q, sizeof(double), g_ctx.stream)) {
    gpu_fail("dot");
    return NAN;
  }
  return r;
}

/* femVecVecMultBig, fem/fem_math.c:1017-1100 */
double femVecVecMultBig(tVector *a, tVector *b)
{
  if (a->rows != b->rows || a->rows <= 0) return 0.0;
  if (a->type != VEC_FULL || b->type != VEC_FULL) { ufem_set_error("VEC_FULL only"); return NAN; }
  if (vec_upload2(a, b) != AF_OK) return NAN;
  return dot_dev(a->rows, g_ctx.d_x, g_ctx.d_y);
}

/* femVecNormBig, fem/fem_math.c:2022-2100 */
double femVecNormBig(tVector *a)
{
  if (a->rows <= 0) return 0.0;
  if (a->type != VEC_FULL) { ufem_set_error("VEC_FULL only"); return NAN; }
  if (vec_upload2(a, NULL) != AF_OK) return NAN;
  return sqrt(dot_dev(a->rows, g_ctx.d_x, g_ctx.d_x));
}

/* femMatNormBig, fem/fem_math.c:1856-1949 */
double femMatNormBig(tMatrix *a)
{
  ugk_state h;
  if (a->type != MAT_SPAR) { ufem_set_error("MAT_SPAR only"); return NAN; }
  if (a->rows <= 0) return 0.0;
  if (ctx_ready() != AF_OK || upload_matrix(a) != AF_OK || ctx_vectors(a->rows, 0) != AF_OK) return NAN;
  memset(&h, 0, sizeof(h));
  if (ugk_h2d(g_ctx.w.st, &h, sizeof(h), g_ctx.stream) ||
      ugk_mat_norm(&g_ctx.K, g_ctx.w.st, g_ctx.w.part, g_ctx.stream) ||
      ugk_d2h(&h, g_ctx.w.st, sizeof(h), g_ctx.stream)) {
    gpu_fail("matrix norm");
    return NAN;
  }
  return h.normA;
}

/* femVecLinCombBig, fem/fem_math.c:1543 */
int femVecLinCombBig(double amult, tVector *a, double bmult, tVector *b, tVector *c)
{
  long n = a->rows;
  if (a->rows != b->rows || b->rows != c->rows) return AF_ERR_SIZ;
  if (a->type != VEC_FULL || b->type != VEC_FULL || c->type != VEC_FULL) return AF_ERR_TYP;
  if (n == 0) return AF_OK;
  if (vec_upload2(a, b) != AF_OK) return AF_ERR;
  if (ugk_axpby(n, amult, g_ctx.d_x, bmult, g_ctx.d_y, g_ctx.d_b, g_ctx.stream) ||
      ugk_d2h(c->data, g_ctx.d_b, sizeof(double) * (size_t)n, g_ctx.stream))
    return gpu_fail("lincomb");
  return AF_OK;
}
