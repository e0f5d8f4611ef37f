/*
 * fem_dropin.c -- the drop-in entry points with uFEM's own signatures
 * (include/ufem_fem_math.h): the solvers of fem/fem_eqs.c and the "Big"
 * sparse kernels of fem/fem_math.c, served by the GPU.
 *
 * A row-slot matrix (fem/fem_math.h:47-57) is read as the reference reads it:
 * row i owns slots frompos[i] .. frompos[i]+defpos[i]-1 and ends at the first
 * slot whose 1-based column is <= 0 (fem/fem_math.c:1424-1433).
 *
 * Device mirror.  Each row-slot matrix handed in gets a device copy in a side
 * table keyed by (data, len, rows, cols) -- the tMatrix struct itself cannot
 * grow.  The copy is SELL-64 of 3x3 node blocks when every row triple
 * 3t, 3t+1, 3t+2 holds the same complete node blocks (a node-major brick
 * matrix, supports included, since femApplyDisp only zeroes values), so the
 * drop-in runs the same kernels as the native path; otherwise scalar SELL-64.
 * Building it sorts each row once and keeps, for every filled slot in row
 * order, the device position of its value.  A later call then only reads
 * the host matrix once: it fingerprints the columns (a changed pattern
 * rebuilds the copy) while packing the values into pinned memory, uploads
 * them and scatters them on the device.  Under ufem_gpu_set_mirror_policy(1)
 * a copy stays valid without that read until ufem_gpu_mark_dirty or
 * ufem_gpu_release is called for the matrix.
 *
 * Matrices and vectors of another type (dense MAT_FULL, sparse vectors) go
 * to the reference's own CPU functions when the program provides them under
 * the names <function>_cpu (the object-file rename of fem_math.o/fem_eqs.o
 * shown in the integration notes); without them the call returns AF_ERR_TYP.
 *
 * Host passes over large matrices run on up to 8 POSIX threads, the
 * reference's own cap (AF_MAX_THREADS, fem/fem.h:64).
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "ufem_fem_math.h"
#include "ufem_internal.h"

/* ------------------------------------------------------------------------ */
/* the reference's CPU functions under their renamed symbols (weak: absent
 * unless the program links them) */
extern int femEqsCGwJ_cpu(tMatrix *a, tVector *b, tVector *x, double eps, long maxIt) __attribute__((weak));
extern int femEqsBiCCSwJ_cpu(tMatrix *a, tVector *b, tVector *x, double eps, long maxIt) __attribute__((weak));
extern int femEqsCGwSSOR_cpu(tMatrix *a, tVector *b, tVector *x, double eps, long maxIt) __attribute__((weak));
extern int femMatVecMultBig_cpu(tMatrix *a, tVector *b, tVector *c) __attribute__((weak));
extern double femVecVecMultBig_cpu(tVector *a, tVector *b) __attribute__((weak));
extern double femVecNormBig_cpu(tVector *a) __attribute__((weak));
extern double femMatNormBig_cpu(tMatrix *a) __attribute__((weak));
extern int femVecLinCombBig_cpu(double amult, tVector *a, double bmult, tVector *b, tVector *c)
    __attribute__((weak));

typedef int (*solver_fn)(tMatrix *, tVector *, tVector *, double, long);

/* the weak symbol if the link resolved it, else a global lookup (a program
 * that loads the CPU functions from a shared object at run time) */
static void *cpu_fn(void *weak, const char *name)
{
  return weak ? weak : dlsym(RTLD_DEFAULT, name);
}
#define CPU_FN(name) cpu_fn((void *)(name##_cpu), #name "_cpu")

static int no_cpu(const char *fn)
{
  ufem_set_error("%s: the GPU serves MAT_SPAR matrices and VEC_FULL vectors; "
                 "no %s_cpu linked for other types", fn, fn);
  return AF_ERR_TYP;
}

/* ------------------------------------------------------------------------ */
static int g_rule = 0;
static int g_policy = 0;
static ufem_solve_info g_last;

void ufem_gpu_set_stop_rule(int rule) { g_rule = rule ? 1 : 0; }
void ufem_gpu_set_mirror_policy(int policy) { g_policy = policy ? 1 : 0; }
void ufem_gpu_last_solve(ufem_solve_info *info)
{
  if (info) *info = g_last;
}

/* UFEM_TRACE set to anything but "" or "0": one line on stderr per solver
 * call (route, size, iterations, status, device time) */
static int trace_on(void)
{
  static int t = -1;
  if (t < 0) {
    const char *e = getenv("UFEM_TRACE");
    t = e && *e && strcmp(e, "0") != 0;
  }
  return t;
}

/* a failure of the device itself is printed as well as returned: the
 * program that called the drop-in may print nothing but its own status */
static int gpu_fail(const char *what)
{
  ufem_set_error("%s: %s", what, ugk_error());
  fprintf(stderr, "[ufem] GPU error: %s\n", ufem_gpu_last_error());
  return AF_ERR;
}

/* ---- host threads -------------------------------------------------------- */
#define PAR_CHUNKS 64
typedef void (*chunk_fn)(void *ctx, int c, long lo, long hi);
typedef struct {
  chunk_fn fn;
  void *ctx;
  long n;
  int next;
} par_job;

static int host_threads(void)
{
  static int t = 0;
  if (!t) {
    const char *e = getenv("UFEM_HOST_THREADS");
    long p = sysconf(_SC_NPROCESSORS_ONLN);
    t = e ? atoi(e) : (int)(p < 1 ? 1 : p);
    if (t < 1) t = 1;
    if (t > 8) t = 8;
  }
  return t;
}

static void *par_worker(void *arg)
{
  par_job *j = (par_job *)arg;
  int c;
  while ((c = __atomic_fetch_add(&j->next, 1, __ATOMIC_RELAXED)) < PAR_CHUNKS)
    j->fn(j->ctx, c, j->n * c / PAR_CHUNKS, j->n * (c + 1) / PAR_CHUNKS);
  return NULL;
}

/* fn over PAR_CHUNKS fixed chunks of [0, n): chunk boundaries do not depend
 * on the thread count, so per-chunk results are reproducible */
static void par_run(long n, chunk_fn fn, void *ctx)
{
  pthread_t th[8];
  int nt = n < 4096 ? 1 : host_threads(), started = 0, i;
  par_job j = {fn, ctx, n, 0};
  for (i = 1; i < nt; i++)
    if (pthread_create(&th[started], NULL, par_worker, &j) == 0) started++;
  par_worker(&j);
  for (i = 0; i < started; i++) pthread_join(th[i], NULL);
}

/* ---- one device context for the drop-in calls, grown on demand ----------- */
typedef struct {
  int ready;
  void *stream, *ev0, *ev1;
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

static int ctx_vectors(long n)
{
  if (g_ctx.n >= n && g_ctx.w.n >= n && g_ctx.w.rhat) return AF_OK;
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

/* ---- the mirror side table ---------------------------------------------- */
#define MIRRORS 8
typedef struct {
  const double *key_data;   /* key: data pointer, len, rows, cols            */
  long key_len, key_rows, key_cols;
  unsigned long long phash; /* fingerprint of the filled columns              */
  long nfill;               /* filled slots                                  */
  int in_use, clean;
  unsigned long tick;       /* last use, for eviction                        */
  ugk_mat K;                /* device copy                                   */
  int *d_map;               /* [nfill]: value position of each filled slot    */
  double *d_stage, *h_stage;/* [nfill] values in host row order               */
  double t_build_ms, t_refresh_ms;
} mirror;

static mirror g_mir[MIRRORS];
static unsigned long g_tick = 0, g_gen = 0;

static void mirror_free(mirror *e)
{
  if (g_ctx.ready) ugk_sync(g_ctx.stream);
  ugk_free(e->K.d_slen); ugk_free(e->K.d_soff); ugk_free(e->K.d_rowlen);
  ugk_free(e->K.d_cols); ugk_free(e->K.d_vals); ugk_free(e->K.d_vals32); ugk_free(e->K.d_corr32); ugk_free(e->K.d_rem16);
  ugk_free(e->d_map); ugk_free(e->d_stage);
  ugk_host_free(e->h_stage);
  /* the SSOR colour order cached for this copy is keyed by its address */
  if (g_ctx.w.ssor && g_ctx.w.ssor->src == &e->K) { ufem_ssor_free(g_ctx.w.ssor); g_ctx.w.ssor = NULL; }
  memset(e, 0, sizeof(*e));
}

static mirror *mirror_find(const tMatrix *a)
{
  int i;
  for (i = 0; i < MIRRORS; i++)
    if (g_mir[i].in_use && g_mir[i].key_data == a->data && g_mir[i].key_len == a->len &&
        g_mir[i].key_rows == a->rows && g_mir[i].key_cols == a->cols)
      return &g_mir[i];
  return NULL;
}

void ufem_gpu_mark_dirty(const tMatrix *a)
{
  mirror *e = a ? mirror_find(a) : NULL;
  if (e) e->clean = 0;
}

void ufem_gpu_release(const tMatrix *a)
{
  mirror *e = a ? mirror_find(a) : NULL;
  if (e) mirror_free(e);
}

/* ---- pass 1: row lengths, column check, fingerprint ----------------------- */
typedef struct {
  const tMatrix *a;
  long *rowlen;
  unsigned long long h[PAR_CHUNKS];
  long fill[PAR_CHUNKS], maxrow[PAR_CHUNKS];
  long bad_row[PAR_CHUNKS];   /* first row with a column out of range, or -1 */
} scan_ctx;

static inline unsigned long long mix64(unsigned long long h, unsigned long long v)
{
  h ^= v + 0x9E3779B97F4A7C15ULL + (h << 6) + (h >> 2);
  h *= 0xBF58476D1CE4E5B9ULL;
  return h ^ (h >> 31);
}

static void scan_chunk(void *vc, int c, long lo, long hi)
{
  scan_ctx *s = (scan_ctx *)vc;
  const tMatrix *a = s->a;
  unsigned long long h = (unsigned long long)c + 1;
  long i, fill = 0, mx = 0;
  s->bad_row[c] = -1;
  for (i = lo; i < hi; i++) {
    const long f = a->frompos[i], cap = a->defpos[i];
    long k;
    if (f < 0 || cap < 0 || f + cap > a->len) { if (s->bad_row[c] < 0) s->bad_row[c] = i; s->rowlen[i] = 0; continue; }
    for (k = 0; k < cap; k++) {
      const long c1 = a->pos[f + k];
      if (c1 <= 0) break;
      if (c1 > a->cols && s->bad_row[c] < 0) s->bad_row[c] = i;
      h = mix64(h, (unsigned long long)c1);
    }
    h = mix64(h, (unsigned long long)k ^ 0x5555555555555555ULL);
    s->rowlen[i] = k;
    fill += k;
    if (k > mx) mx = k;
  }
  s->h[c] = h;
  s->fill[c] = fill;
  s->maxrow[c] = mx;
}

/* ---- pass 3: pack the filled values in row order -------------------------- */
typedef struct {
  const tMatrix *a;
  const long *rowlen, *foff;
  double *stage;
} pack_ctx;

static void pack_chunk(void *vc, int c, long lo, long hi)
{
  pack_ctx *p = (pack_ctx *)vc;
  long i;
  (void)c;
  for (i = lo; i < hi; i++)
    if (p->rowlen[i] > 0)
      memcpy(p->stage + p->foff[i], p->a->data + p->a->frompos[i], sizeof(double) * (size_t)p->rowlen[i]);
}

/* ---- build: 3x3 node blocks or scalar rows -------------------------------- */
typedef struct {
  const tMatrix *a;
  const long *rowlen, *foff;
  long maxrow;
  int bs;
  /* bs 3: sorted node blocks of each row triple at blk[foff[3t] / 9] */
  int *blk;
  /* SELL arrays */
  const long *soff;
  int *cols;
  int *map;
  int fail[PAR_CHUNKS];     /* 1: not a block matrix; 2: duplicate column */
} build_ctx;

static void isort_int(int *v, long n)
{
  long i, j;
  for (i = 1; i < n; i++) {
    int x = v[i];
    for (j = i; j > 0 && v[j - 1] > x; j--) v[j] = v[j - 1];
    v[j] = x;
  }
}

static int cmp_int(const void *x, const void *y)
{
  int a = *(const int *)x, b = *(const int *)y;
  return (a > b) - (a < b);
}

static void sort_int(int *v, long n)
{
  if (n <= 48) isort_int(v, n);
  else qsort(v, (size_t)n, sizeof(int), cmp_int);
}

static long find_int(const int *v, long n, int x)
{
  long lo = 0, hi = n;
  while (lo < hi) {
    long mid = (lo + hi) >> 1;
    if (v[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return (lo < n && v[lo] == x) ? lo : -1;
}

/* row triples t in [lo, hi): the node blocks of row 3t, and every entry of
 * the three rows inside them exactly once */
static void detect3_chunk(void *vc, int c, long lo, long hi)
{
  build_ctx *b = (build_ctx *)vc;
  const tMatrix *a = b->a;
  int *tmp = (int *)malloc(sizeof(int) * (size_t)(b->maxrow + 1));
  unsigned char *seen = (unsigned char *)malloc((size_t)b->maxrow + 1);
  long t;
  b->fail[c] = 0;
  if (!tmp || !seen) { b->fail[c] = 1; free(tmp); free(seen); return; }
  for (t = lo; t < hi && !b->fail[c]; t++) {
    const long r0 = 3 * t, L = b->rowlen[r0];
    long k, nb = 0, r;
    int *blk = b->blk + b->foff[r0] / 9;
    if (L % 3 || b->rowlen[r0 + 1] != L || b->rowlen[r0 + 2] != L || b->foff[r0] % 9) { b->fail[c] = 1; break; }
    for (k = 0; k < L; k++) tmp[k] = (int)((a->pos[a->frompos[r0] + k] - 1) / 3);
    sort_int(tmp, L);
    for (k = 0; k < L; k++)
      if (k == 0 || tmp[k] != tmp[k - 1]) tmp[nb++] = tmp[k];
    if (3 * nb != L) { b->fail[c] = 1; break; }
    memcpy(blk, tmp, sizeof(int) * (size_t)nb);
    for (r = 0; r < 3 && !b->fail[c]; r++) {
      memset(seen, 0, (size_t)L);
      for (k = 0; k < L; k++) {
        const long col = a->pos[a->frompos[r0 + r] + k] - 1;
        const long q = find_int(blk, nb, (int)(col / 3));
        if (q < 0) { b->fail[c] = 1; break; }
        if (seen[3 * q + col % 3]) { b->fail[c] = 2; break; }
        seen[3 * q + col % 3] = 1;
      }
    }
  }
  free(tmp);
  free(seen);
}

/* SELL columns and the value map, block rows [lo, hi) */
static void fill3_chunk(void *vc, int c, long lo, long hi)
{
  build_ctx *b = (build_ctx *)vc;
  const tMatrix *a = b->a;
  long t;
  (void)c;
  for (t = lo; t < hi; t++) {
    const long r0 = 3 * t, L = b->rowlen[r0], nb = L / 3, S0 = b->soff[t >> 6];
    const int l = (int)(t & 63);
    const int *blk = b->blk + b->foff[r0] / 9;
    long k, r;
    for (k = 0; k < nb; k++) b->cols[(S0 + k) * 64 + l] = blk[k];
    for (r = 0; r < 3; r++)
      for (k = 0; k < L; k++) {
        const long col = a->pos[a->frompos[r0 + r] + k] - 1;
        const long q = find_int(blk, nb, (int)(col / 3));
        b->map[b->foff[r0 + r] + k] = (int)ugk_vpos3(S0 + q, (int)(3 * r + col % 3), l);
      }
  }
}

typedef struct { int col, idx; } colidx;
static int cmp_colidx(const void *x, const void *y)
{
  int a = ((const colidx *)x)->col, b = ((const colidx *)y)->col;
  return (a > b) - (a < b);
}

static void fill1_chunk(void *vc, int c, long lo, long hi)
{
  build_ctx *b = (build_ctx *)vc;
  const tMatrix *a = b->a;
  colidx *tmp = (colidx *)malloc(sizeof(colidx) * (size_t)(b->maxrow + 1));
  long i;
  b->fail[c] = 0;
  if (!tmp) { b->fail[c] = 1; return; }
  for (i = lo; i < hi; i++) {
    const long L = b->rowlen[i], S0 = b->soff[i >> 6];
    const int l = (int)(i & 63);
    long k;
    for (k = 0; k < L; k++) {
      tmp[k].col = (int)(a->pos[a->frompos[i] + k] - 1);
      tmp[k].idx = (int)k;
    }
    qsort(tmp, (size_t)L, sizeof(colidx), cmp_colidx);
    for (k = 0; k < L; k++) {
      if (k > 0 && tmp[k].col == tmp[k - 1].col) { b->fail[c] = 2; break; }
      b->cols[(S0 + k) * 64 + l] = tmp[k].col;
      b->map[b->foff[i] + tmp[k].idx] = (int)((S0 + k) * 64 + l);
    }
  }
  free(tmp);
}

static double now_ms(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return 1e3 * (double)ts.tv_sec + 1e-6 * (double)ts.tv_nsec;
}

/* device copy of the pattern (and the value map) from the scanned matrix */
static int mirror_build(mirror *e, const tMatrix *a, const long *rowlen, const long *foff, long nfill, long maxrow)
{
  build_ctx b;
  const long n = a->rows;
  long nbr, nsl, s, total = 0, nnzb = 0, i;
  int bs = 1, rv = AF_ERR_MEM, c;
  int *blen = NULL, *slen = NULL, *hcols = NULL, *hmap = NULL, *blk = NULL;
  long *soff = NULL;
  double t0 = now_ms();

  memset(&b, 0, sizeof(b));
  b.a = a;
  b.rowlen = rowlen;
  b.foff = foff;
  b.maxrow = maxrow;
  /* 3x3 node blocks? */
  if (a->rows == a->cols && n % 3 == 0 && n >= 3 && maxrow > 0) {
    blk = (int *)malloc(sizeof(int) * (size_t)(nfill / 9 + 1));
    if (!blk) goto out;
    b.blk = blk;
    par_run(n / 3, detect3_chunk, &b);
    bs = 3;
    for (c = 0; c < PAR_CHUNKS; c++) {
      if (b.fail[c] == 2) { ufem_set_error("duplicate column in a row"); rv = AF_ERR_VAL; goto out; }
      if (b.fail[c]) bs = 1;
    }
  }
  nbr = n / bs;
  nsl = (nbr + 63) / 64;
  if (nsl > 4L * UGK_RED_MAX_BLOCKS) { ufem_set_error("matrix too large (%ld rows)", n); rv = AF_ERR_BIG; goto out; }
  blen = (int *)malloc(sizeof(int) * (size_t)(nbr + 1));
  slen = (int *)calloc((size_t)nsl + 1, sizeof(int));
  soff = (long *)calloc((size_t)nsl + 1, sizeof(long));
  if (!blen || !slen || !soff) goto out;
  for (i = 0; i < nbr; i++) {
    blen[i] = (int)(bs == 3 ? rowlen[3 * i] / 3 : rowlen[i]);
    nnzb += blen[i];
    if (blen[i] > slen[i >> 6]) slen[i >> 6] = blen[i];
  }
  for (s = 0; s < nsl; s++) {
    soff[s] = total;
    total += slen[s];
  }
  if ((double)total * 64.0 * bs * bs >= 2147483647.0) {
    ufem_set_error("matrix too large for the value map (%ld slots)", total);
    rv = AF_ERR_BIG;
    goto out;
  }
  hcols = (int *)malloc(sizeof(int) * (size_t)(total * 64 + 1));
  hmap = (int *)malloc(sizeof(int) * (size_t)(nfill + 1));
  if (!hcols || !hmap) goto out;
  /* padding slots: the row itself (value 0), a column inside the matrix */
  for (s = 0; s < nsl; s++) {
    long k;
    int l;
    for (l = 0; l < 64; l++) {
      const long I = s * 64 + l;
      const long ncb = a->cols / bs;
      const int pad = (int)(I < ncb ? I : (ncb > 0 ? ncb - 1 : 0));
      for (k = 0; k < slen[s]; k++) hcols[(soff[s] + k) * 64 + l] = pad;
    }
  }
  b.bs = bs;
  b.soff = soff;
  b.cols = hcols;
  b.map = hmap;
  memset(b.fail, 0, sizeof(b.fail));
  if (bs == 3) par_run(nbr, fill3_chunk, &b);
  else par_run(nbr, fill1_chunk, &b);
  for (c = 0; c < PAR_CHUNKS; c++) {
    if (b.fail[c] == 2) { ufem_set_error("duplicate column in a row"); rv = AF_ERR_VAL; goto out; }
    if (b.fail[c]) goto out;
  }
  /* device arrays (values zeroed: padding stays 0) */
  mirror_free(e);
  e->K.bs = bs;
  e->K.n_brows = nbr;
  e->K.n_slices = nsl;
  e->K.total_slots = total;
  e->K.nnzb = nnzb;
  e->K.gen = ++g_gen;
  e->K.d_slen = (int *)ugk_malloc(sizeof(int) * (size_t)(nsl + 1));
  e->K.d_soff = (long *)ugk_malloc(sizeof(long) * (size_t)(nsl + 1));
  e->K.d_rowlen = (int *)ugk_malloc(sizeof(int) * (size_t)(nbr + 1));
  e->K.d_cols = (int *)ugk_malloc(sizeof(int) * (size_t)(total * 64 + 64));
  e->K.d_vals = (double *)ugk_malloc(sizeof(double) * (size_t)(total * 64 * bs * bs + 64));
  e->d_map = (int *)ugk_malloc(sizeof(int) * (size_t)(nfill + 1));
  e->d_stage = (double *)ugk_malloc(sizeof(double) * (size_t)(nfill + 1));
  e->h_stage = (double *)ugk_host_alloc(sizeof(double) * (size_t)(nfill + 1));
  if (!e->K.d_slen || !e->K.d_soff || !e->K.d_rowlen || !e->K.d_cols || !e->K.d_vals || !e->d_map ||
      !e->d_stage || !e->h_stage) {
    mirror_free(e);
    rv = gpu_fail("device copy of the matrix");
    goto out;
  }
  if (ugk_h2d(e->K.d_slen, slen, sizeof(int) * (size_t)nsl, g_ctx.stream) ||
      ugk_h2d(e->K.d_soff, soff, sizeof(long) * (size_t)nsl, g_ctx.stream) ||
      ugk_h2d(e->K.d_rowlen, blen, sizeof(int) * (size_t)nbr, g_ctx.stream) ||
      ugk_h2d(e->K.d_cols, hcols, sizeof(int) * (size_t)total * 64, g_ctx.stream) ||
      ugk_h2d(e->d_map, hmap, sizeof(int) * (size_t)nfill, g_ctx.stream) || ugk_sync(g_ctx.stream)) {
    mirror_free(e);
    rv = gpu_fail("matrix upload");
    goto out;
  }
  e->nfill = nfill;
  e->in_use = 1;
  e->t_build_ms = now_ms() - t0;
  rv = AF_OK;
out:
  free(blk); free(blen); free(slen); free(soff); free(hcols); free(hmap);
  return rv;
}

/* the device copy of a row-slot matrix, built or refreshed as the policy
 * asks; NULL on error (code in *rv) */
static mirror *mirror_get(const tMatrix *a, int *rv)
{
  mirror *e = mirror_find(a);
  scan_ctx *sc = NULL;
  long *foff = NULL, nfill = 0, maxrow = 0, i;
  unsigned long long h = 0;
  double t0;
  int c;
  *rv = AF_OK;
  if (e && g_policy == 1 && e->clean) {
    e->tick = ++g_tick;
    e->t_refresh_ms = 0.0;
    return e;
  }
  t0 = now_ms();
  sc = (scan_ctx *)calloc(1, sizeof(scan_ctx));
  foff = (long *)malloc(sizeof(long) * (size_t)(a->rows + 1));
  if (!sc || !foff || !(sc->rowlen = (long *)malloc(sizeof(long) * (size_t)(a->rows + 1)))) {
    *rv = AF_ERR_MEM;
    goto fail;
  }
  sc->a = a;
  par_run(a->rows, scan_chunk, sc);
  for (c = 0; c < PAR_CHUNKS; c++) {
    if (sc->bad_row[c] >= 0) {
      ufem_set_error("row %ld: slots outside the matrix or a column beyond %ld", sc->bad_row[c] + 1, a->cols);
      *rv = AF_ERR_SIZ;
      goto fail;
    }
    h = mix64(h, sc->h[c]);
    nfill += sc->fill[c];
    if (sc->maxrow[c] > maxrow) maxrow = sc->maxrow[c];
  }
  foff[0] = 0;
  for (i = 0; i < a->rows; i++) foff[i + 1] = foff[i] + sc->rowlen[i];
  if (!e || e->phash != h || e->nfill != nfill) {
    if (!e) {   /* a free slot, else the least recently used one */
      int k, pick = 0;
      for (k = 0; k < MIRRORS; k++) {
        if (!g_mir[k].in_use) { pick = k; break; }
        if (g_mir[k].tick < g_mir[pick].tick) pick = k;
      }
      e = &g_mir[pick];
      mirror_free(e);
    }
    if ((*rv = mirror_build(e, a, sc->rowlen, foff, nfill, maxrow)) != AF_OK) { mirror_free(e); goto fail; }
    e->key_data = a->data;
    e->key_len = a->len;
    e->key_rows = a->rows;
    e->key_cols = a->cols;
    e->phash = h;
  }
  {
    pack_ctx p = {a, sc->rowlen, foff, e->h_stage};
    par_run(a->rows, pack_chunk, &p);
  }
  if (ugk_h2d(e->d_stage, e->h_stage, sizeof(double) * (size_t)nfill, g_ctx.stream) ||
      ugk_scatter_vals(nfill, e->d_map, e->d_stage, e->K.d_vals, g_ctx.stream) || ugk_sync(g_ctx.stream)) {
    *rv = gpu_fail("matrix values");
    mirror_free(e);
    goto fail;
  }
  e->clean = 1;
  e->tick = ++g_tick;
  e->t_refresh_ms = now_ms() - t0;
  free(sc->rowlen); free(sc); free(foff);
  return e;
fail:
  if (sc) free(sc->rowlen);
  free(sc); free(foff);
  return NULL;
}

/* the device copy's shape and host times, for benchmarks and tests:
 * info[0] block size, info[1] rows, info[2] filled slots, info[3] last build
 * ms (host + upload), info[4] last refresh ms (read, pack, upload, scatter;
 * 0 when a clean copy was used) */
int ufem_gpu_mirror_info(const tMatrix *a, double *info)
{
  mirror *e = a ? mirror_find(a) : NULL;
  if (!e) { ufem_set_error("no device copy of this matrix"); return AF_ERR_VAL; }
  info[0] = e->K.bs;
  info[1] = (double)e->K.n_brows * e->K.bs;
  info[2] = (double)e->nfill;
  info[3] = e->t_build_ms;
  info[4] = e->t_refresh_ms;
  return AF_OK;
}

/* ---- solvers ---------------------------------------------------------- */
static int served(const tMatrix *a, const tVector *b, const tVector *x)
{
  return a->type == MAT_SPAR && b->type == VEC_FULL && x->type == VEC_FULL;
}

static int dropin_solve_(tMatrix *a, tVector *b, tVector *x, double eps, long maxIt, int solver);

static int dropin_solve(tMatrix *a, tVector *b, tVector *x, double eps, long maxIt, int solver,
                        const char *name)
{
  int rv = dropin_solve_(a, b, x, eps, maxIt, solver);
  if (trace_on())
    fprintf(stderr, "[ufem] %s on the GPU: n %ld, %ld iterations, rv %d, %.2f ms\n", name, a->rows,
            g_last.iters, rv, g_last.solve_ms);
  return rv;
}

static solver_fn cpu_solver(solver_fn f, const char *name)
{
  if (f && trace_on()) fprintf(stderr, "[ufem] %s on the CPU (%s_cpu)\n", name, name);
  return f;
}

static int dropin_solve_(tMatrix *a, tVector *b, tVector *x, double eps, long maxIt, int solver)
{
  ufem_solve_opts o;
  mirror *e;
  int rv;
  long n = a->rows, i;
  double bb = 0.0;
  memset(&g_last, 0, sizeof(g_last));
  /* the reference's order: sizes, then a zero load returns x = 0 before
     anything else is looked at (fem/fem_eqs.c:103-116) */
  if (a->cols != x->rows || x->rows != b->rows) return AF_ERR_SIZ;
  if (b->type == VEC_FULL) {
    for (i = 0; i < b->rows; i++) bb += b->data[i] * b->data[i];
    if (!(bb > 0.0)) {
      if (x->type == VEC_FULL) memset(x->data, 0, sizeof(double) * (size_t)x->len);
      else for (i = 0; i < x->len; i++) x->data[i] = 0.0;
      return AF_OK;
    }
  }
  if (a->rows != b->rows) return AF_ERR_SIZ;
  if (n == 0) return AF_OK;
  if ((rv = ctx_ready()) != AF_OK) return rv;
  if (!(e = mirror_get(a, &rv))) return rv;
  if ((rv = ctx_vectors(n)) != AF_OK) return rv;
  if (ugk_h2d(g_ctx.d_b, b->data, sizeof(double) * (size_t)n, g_ctx.stream) ||
      ugk_h2d(g_ctx.d_x, x->data, sizeof(double) * (size_t)n, g_ctx.stream))
    return gpu_fail("vector upload");
  memset(&o, 0, sizeof(o));
  o.solver = solver;
  o.rule = g_rule;
  o.eps = eps;
  o.maxit = maxIt > 0 ? maxIt : -1;   /* maxIt <= 0: the loop runs zero times */
  o.exec = UFEM_EXEC_STREAMED;
  rv = ufem_solve_device(&e->K, n, g_ctx.d_b, g_ctx.d_x, &g_ctx.w, g_ctx.stream, g_ctx.ev0, g_ctx.ev1,
                         &o, &g_last);
  if (rv == AF_ERR_ZER || rv == AF_ERR_MEM) return rv;
  if (ugk_d2h(x->data, g_ctx.d_x, sizeof(double) * (size_t)n, g_ctx.stream)) return gpu_fail("x download");
  return rv;
}

/* femEqsCGwJ, fem/fem_eqs.c:88-247 */
int femEqsCGwJ(tMatrix *a, tVector *b, tVector *x, double eps, long maxIt)
{
  if (!served(a, b, x)) {
    solver_fn f = cpu_solver((solver_fn)CPU_FN(femEqsCGwJ), "femEqsCGwJ");
    return f ? f(a, b, x, eps, maxIt) : no_cpu("femEqsCGwJ");
  }
  return dropin_solve(a, b, x, eps, maxIt, UFEM_SOLVER_PCG, "femEqsCGwJ");
}

/* femEqsCGwSSOR, fem/fem_eqs.c:999-1188 (colour-ordered sweeps, see header) */
int femEqsCGwSSOR(tMatrix *a, tVector *b, tVector *x, double eps, long maxIt)
{
  if (!served(a, b, x)) {
    solver_fn f = cpu_solver((solver_fn)CPU_FN(femEqsCGwSSOR), "femEqsCGwSSOR");
    return f ? f(a, b, x, eps, maxIt) : no_cpu("femEqsCGwSSOR");
  }
  return dropin_solve(a, b, x, eps, maxIt, UFEM_SOLVER_SSOR, "femEqsCGwSSOR");
}

/* femEqsBiCCSwJ, fem/fem_eqs.c:262-468 (true-residual test, see header) */
int femEqsBiCCSwJ(tMatrix *a, tVector *b, tVector *x, double eps, long maxIt)
{
  int rv;
  if (!served(a, b, x)) {
    solver_fn f = cpu_solver((solver_fn)CPU_FN(femEqsBiCCSwJ), "femEqsBiCCSwJ");
    return f ? f(a, b, x, eps, maxIt) : no_cpu("femEqsBiCCSwJ");
  }
  rv = dropin_solve(a, b, x, eps, maxIt, UFEM_SOLVER_BICGSTAB, "femEqsBiCCSwJ");
  /* the reference reports an unconverged run on msgout and returns AF_OK
     with the last x (fem/fem_eqs.c:441-467); the status stays in
     ufem_gpu_last_solve */
  if (rv == AF_ERR && g_last.status == AF_ERR) {
    fprintf(stderr, "[ufem] femEqsBiCCSwJ: unconverged solution (true relative residual %.3e after %ld "
            "iterations)\n", g_last.true_relres, g_last.iters);
    rv = AF_OK;
  }
  return rv;
}

/* ---- the "Big" kernels ------------------------------------------------- */
/* femMatVecMultBig, fem/fem_math.c:1350-1451 */
int femMatVecMultBig(tMatrix *a, tVector *b, tVector *c)
{
  int rv;
  long n;
  mirror *e;
  if (a->cols != b->rows || c->rows != a->rows) return AF_ERR_SIZ;
  if (a->type != MAT_SPAR || b->type != VEC_FULL || c->type != VEC_FULL) {
    int (*f)(tMatrix *, tVector *, tVector *) = (int (*)(tMatrix *, tVector *, tVector *))CPU_FN(femMatVecMultBig);
    return f ? f(a, b, c) : no_cpu("femMatVecMultBig");
  }
  n = a->rows;
  if (n == 0) return AF_OK;
  if ((rv = ctx_ready()) != AF_OK) return rv;
  if (!(e = mirror_get(a, &rv))) return rv;
  if ((rv = ctx_vectors(n > b->rows ? n : b->rows)) != AF_OK) return rv;
  if (ugk_h2d(g_ctx.d_x, b->data, sizeof(double) * (size_t)b->rows, g_ctx.stream) ||
      ugk_spmv(&e->K, g_ctx.d_x, g_ctx.d_y, g_ctx.stream) ||
      ugk_d2h(c->data, g_ctx.d_y, sizeof(double) * (size_t)n, g_ctx.stream))
    return gpu_fail("spmv");
  return AF_OK;
}

static int vec_upload2(tVector *a, tVector *b)
{
  int rv;
  long n = a->rows;
  if ((rv = ctx_ready()) != AF_OK) return rv;
  if ((rv = ctx_vectors(n)) != AF_OK) return rv;
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
  if (a->type != VEC_FULL || b->type != VEC_FULL) {
    double (*f)(tVector *, tVector *) = (double (*)(tVector *, tVector *))CPU_FN(femVecVecMultBig);
    if (f) return f(a, b);
    no_cpu("femVecVecMultBig");
    return NAN;
  }
  if (a->rows != b->rows || a->rows <= 0) return 0.0;
  if (vec_upload2(a, b) != AF_OK) return NAN;
  return dot_dev(a->rows, g_ctx.d_x, g_ctx.d_y);
}

/* femVecNormBig, fem/fem_math.c:2022-2100 */
double femVecNormBig(tVector *a)
{
  if (a->type != VEC_FULL) {
    double (*f)(tVector *) = (double (*)(tVector *))CPU_FN(femVecNormBig);
    if (f) return f(a);
    no_cpu("femVecNormBig");
    return NAN;
  }
  if (a->rows <= 0) return 0.0;
  if (vec_upload2(a, NULL) != AF_OK) return NAN;
  return sqrt(dot_dev(a->rows, g_ctx.d_x, g_ctx.d_x));
}

/* femMatNormBig, fem/fem_math.c:1856-1949 */
double femMatNormBig(tMatrix *a)
{
  ugk_state h;
  mirror *e;
  int rv;
  if (a->type != MAT_SPAR) {
    double (*f)(tMatrix *) = (double (*)(tMatrix *))CPU_FN(femMatNormBig);
    if (f) return f(a);
    no_cpu("femMatNormBig");
    return NAN;
  }
  if (a->rows <= 0) return 0.0;
  if (ctx_ready() != AF_OK || !(e = mirror_get(a, &rv)) || ctx_vectors(a->rows) != AF_OK) return NAN;
  memset(&h, 0, sizeof(h));
  if (ugk_h2d(g_ctx.w.st, &h, sizeof(h), g_ctx.stream) ||
      ugk_mat_norm(&e->K, g_ctx.w.st, g_ctx.w.part, g_ctx.stream) ||
      ugk_d2h(&h, g_ctx.w.st, sizeof(h), g_ctx.stream)) {
    gpu_fail("matrix norm");
    return NAN;
  }
  return h.normA;
}

/* femVecLinCombBig, fem/fem_math.c:1543-1612 */
int femVecLinCombBig(double amult, tVector *a, double bmult, tVector *b, tVector *c)
{
  long n = a->rows;
  if (a->type != VEC_FULL || b->type != VEC_FULL || c->type != VEC_FULL) {
    int (*f)(double, tVector *, double, tVector *, tVector *) =
        (int (*)(double, tVector *, double, tVector *, tVector *))CPU_FN(femVecLinCombBig);
    return f ? f(amult, a, bmult, b, c) : no_cpu("femVecLinCombBig");
  }
  if (a->rows != b->rows || b->rows != c->rows) return AF_ERR_SIZ;
  if (n == 0) return AF_OK;
  if (vec_upload2(a, b) != AF_OK) return AF_ERR;
  if (ugk_axpby(n, amult, g_ctx.d_x, bmult, g_ctx.d_y, g_ctx.d_b, g_ctx.stream) ||
      ugk_d2h(c->data, g_ctx.d_b, sizeof(double) * (size_t)n, g_ctx.stream))
    return gpu_fail("lincomb");
  return AF_OK;
}

/* femMatSetZeroBig, fem/fem_math.c:483-523: every stored value (len of them)
 * to 0, the slots' columns kept.  A device copy of the matrix is cleared
 * with it and stays valid. */
void femMatSetZeroBig(tMatrix *a)
{
  mirror *e = mirror_find(a);
  if (a->len > 0 && a->data) memset(a->data, 0, sizeof(double) * (size_t)a->len);
  if (e) {
    if (ugk_zero(e->K.d_vals, sizeof(double) * (size_t)e->K.total_slots * 64 * e->K.bs * e->K.bs, g_ctx.stream) ||
        ugk_sync(g_ctx.stream))
      mirror_free(e);
  }
}

/* femVecSetZeroBig, fem/fem_math.c:920-958: every stored value (len) to 0 */
void femVecSetZeroBig(tVector *a)
{
  if (a->len > 0 && a->data) memset(a->data, 0, sizeof(double) * (size_t)a->len);
}
