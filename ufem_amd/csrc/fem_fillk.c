/*
 * fem_fillk.c -- fem_fill_K(AF_NO) (fem/fem_sol.c:897-989) on the GPU for a
 * model whose elements are all 8-node bricks (type 9) of material type 1:
 * the element matrices (e009_stiff Mode=AF_NO, fem/fem_e009.c:261-635) are
 * computed and summed on the device with the colour-ordered scatter of the
 * native path, and the sum is added into the reference's own row-slot K
 * through its DOF map nDOFfld, as femLocKM_e -> femMatPutAdd(FEM_ADD) does
 * (fem/fem_mloc.c:121-221, fem/fem_math.c:185-235): an existing column of a
 * row takes the value added to it, a new one goes to the row's first empty
 * slot.  Rows and columns whose DOF the map leaves at 0 (a support removed
 * by -fbc, fem/fem_sol.c:1280-1296, or a DOF of no element) are skipped, as
 * the reference's localisation skips them (fem/fem_mloc.c:185-192).
 *
 * The device system (pattern, colouring) is kept between calls while the
 * connectivity stays the same, so a refill after new moduli or coordinates
 * only uploads those.  The device copy the drop-in solvers keep of K is
 * marked stale afterwards.
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "ufem_fem_math.h"
#include "ufem_internal.h"

void ufem_gpu_mark_dirty(const tMatrix *a);   /* fem_dropin.c */

static struct {
  ufem_sys *sys;
  long n_used, n_elems;
  long *enodes;             /* renumbered connectivity of the cached system */
  double t[3];              /* set-up, device fill, write-back (ms)         */
} g_fill;

static double now_ms(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return 1e3 * (double)ts.tv_sec + 1e-6 * (double)ts.tv_nsec;
}

void ufem_gpu_fill_K_timings(double t[3])
{
  memcpy(t, g_fill.t, sizeof(g_fill.t));
}

void ufem_gpu_fill_K_release(void)
{
  ufem_sys_destroy(g_fill.sys);
  free(g_fill.enodes);
  memset(&g_fill, 0, sizeof(g_fill));
}

/* the device matrix downloaded, and the row-slot target */
typedef struct {
  const ufem_sys *s;
  const int *cols, *rowlen;
  const long *soff;
  const double *vals;
  const long *used;         /* device node -> model node */
  const long *map;          /* nDOFfld */
  long known;
  tMatrix *K;
  int err[64];              /* per chunk: 0, AF_ERR_BIG (row full) or AF_ERR_SIZ */
  long err_row[64];
} wb_ctx;

typedef struct { long col, slot; } cslot;
static int cmp_cslot(const void *x, const void *y)
{
  long a = ((const cslot *)x)->col, b = ((const cslot *)y)->col;
  return (a > b) - (a < b);
}

/* add the entries (C[k], V[k]) into row R (1-based) of the row-slot K */
static int add_row(tMatrix *K, long R, const long *C, const double *V, long cnt, cslot *tmp)
{
  const long f = K->frompos[R - 1], cap = K->defpos[R - 1];
  long L = 0, k;
  while (L < cap && K->pos[f + L] > 0) L++;
  if (L == 0) {              /* an empty row (first fill): in order */
    if (cnt > cap) return AF_ERR_BIG;
    for (k = 0; k < cnt; k++) {
      K->pos[f + k] = C[k];
      K->data[f + k] = V[k];
    }
    return AF_OK;
  }
  for (k = 0; k < L; k++) {
    tmp[k].col = K->pos[f + k];
    tmp[k].slot = f + k;
  }
  qsort(tmp, (size_t)L, sizeof(cslot), cmp_cslot);
  {
    const long L0 = L;
    for (k = 0; k < cnt; k++) {
      long lo = 0, hi = L0;
      while (lo < hi) {
        long mid = (lo + hi) >> 1;
        if (tmp[mid].col < C[k]) lo = mid + 1;
        else hi = mid;
      }
      if (lo < L0 && tmp[lo].col == C[k]) K->data[tmp[lo].slot] += V[k];
      else {
        if (L >= cap) return AF_ERR_BIG;
        K->pos[f + L] = C[k];
        K->data[f + L] = V[k];
        L++;
      }
    }
  }
  return AF_OK;
}

static void wb_chunk(wb_ctx *w, int c, long lo, long hi)
{
  const long mb = (long)w->s->maxb * 3;
  long *C = (long *)malloc(sizeof(long) * (size_t)(mb + 1));
  double *V = (double *)malloc(sizeof(double) * (size_t)(mb + 1));
  cslot *tmp = NULL;
  long I, tmp_cap = 0;
  w->err[c] = 0;
  if (!C || !V) { w->err[c] = AF_ERR_MEM; goto out; }
  for (I = lo; I < hi && !w->err[c]; I++) {
    const long p = w->used[I], so = w->soff[I >> 6];
    const int ln = (int)(I & 63);
    int r;
    for (r = 0; r < 3; r++) {
      const long R = w->map[p * w->known + r];
      long cnt = 0, k, cap;
      if (R <= 0) continue;
      if (R > w->K->rows) { w->err[c] = AF_ERR_SIZ; w->err_row[c] = R; break; }
      for (k = 0; k < w->rowlen[I]; k++) {
        const long J = w->cols[(so + k) * 64 + ln], pj = w->used[J];
        int q;
        for (q = 0; q < 3; q++) {
          const long Cq = w->map[pj * w->known + q];
          if (Cq <= 0) continue;
          C[cnt] = Cq;
          V[cnt] = w->vals[ugk_vpos3(so + k, 3 * r + q, ln)];
          cnt++;
        }
      }
      cap = w->K->defpos[R - 1];
      if (cap > tmp_cap) {   /* room to sort the row's filled slots */
        free(tmp);
        tmp_cap = cap;
        tmp = (cslot *)malloc(sizeof(cslot) * (size_t)tmp_cap);
        if (!tmp) { w->err[c] = AF_ERR_MEM; break; }
      }
      if ((w->err[c] = add_row(w->K, R, C, V, cnt, tmp)) != AF_OK) { w->err_row[c] = R; break; }
    }
  }
out:
  free(C); free(V); free(tmp);
}

typedef struct { wb_ctx *w; long n; int next; } wb_job;
static void *wb_worker(void *arg)
{
  wb_job *j = (wb_job *)arg;
  int c;
  while ((c = __atomic_fetch_add(&j->next, 1, __ATOMIC_RELAXED)) < 64)
    wb_chunk(j->w, c, j->n * c / 64, j->n * (c + 1) / 64);
  return NULL;
}

int ufem_gpu_fill_K(tMatrix *K, const ufem_mesh *m, const long *dof_map, long known_dofs, int npts)
{
  long *newid = NULL, *used = NULL, *en = NULL, i, nu = 0;
  double *x = NULL, *y = NULL, *z = NULL;
  int *hc = NULL, *hl = NULL, rv = AF_ERR_MEM, c;
  long *ho = NULL;
  double *hv = NULL, t0 = now_ms(), t1, t2;
  ufem_mesh dm;
  ufem_sys *s;

  if (!K || !m || !dof_map || K->type != MAT_SPAR || known_dofs < 3 || (npts != 2 && npts != 3)) {
    ufem_set_error("fill_K: a MAT_SPAR matrix, a mesh, the DOF map (>= 3 DOFs per node) and 2 or 3 points");
    return AF_ERR_VAL;
  }
  if (m->n_elems < 1 || m->n_nodes < 1) { ufem_set_error("fill_K: empty mesh"); return AF_ERR_SIZ; }
  /* the nodes elements use, in model order (a node of no element has no DOF) */
  newid = (long *)malloc(sizeof(long) * (size_t)m->n_nodes);
  used = (long *)malloc(sizeof(long) * (size_t)m->n_nodes);
  en = (long *)malloc(sizeof(long) * (size_t)m->n_elems * 8);
  if (!newid || !used || !en) goto out;
  for (i = 0; i < m->n_nodes; i++) newid[i] = -1;
  for (i = 0; i < m->n_elems * 8; i++) {
    if (m->enodes[i] < 0 || m->enodes[i] >= m->n_nodes) {
      ufem_set_error("fill_K: element %ld names node position %ld", i / 8, m->enodes[i]);
      rv = AF_ERR_VAL;
      goto out;
    }
    newid[m->enodes[i]] = 0;
  }
  for (i = 0; i < m->n_nodes; i++)
    if (newid[i] == 0) { used[nu] = i; newid[i] = nu++; }
  for (i = 0; i < m->n_elems * 8; i++) en[i] = newid[m->enodes[i]];
  x = (double *)malloc(sizeof(double) * (size_t)nu);
  y = (double *)malloc(sizeof(double) * (size_t)nu);
  z = (double *)malloc(sizeof(double) * (size_t)nu);
  if (!x || !y || !z) goto out;
  for (i = 0; i < nu; i++) {
    x[i] = m->x[used[i]];
    y[i] = m->y[used[i]];
    z[i] = m->z[used[i]];
  }
  memset(&dm, 0, sizeof(dm));
  dm.n_nodes = nu;
  dm.x = x;
  dm.y = y;
  dm.z = z;
  dm.n_elems = m->n_elems;
  dm.enodes = en;
  dm.E = m->E;
  dm.nu = m->nu;
  /* the cached device system if the connectivity is the same */
  if (g_fill.sys && (g_fill.n_used != nu || g_fill.n_elems != m->n_elems ||
                     memcmp(g_fill.enodes, en, sizeof(long) * (size_t)m->n_elems * 8)))
    ufem_gpu_fill_K_release();
  if (g_fill.sys) {
    if ((rv = ufem_sys_update(g_fill.sys, &dm)) != AF_OK) goto out;
  } else {
    if (!(g_fill.sys = ufem_sys_create(&dm))) { rv = AF_ERR; goto out; }
    g_fill.n_used = nu;
    g_fill.n_elems = m->n_elems;
    g_fill.enodes = en;
    en = NULL;
  }
  s = g_fill.sys;
  if ((rv = ufem_sys_fill_K(s, npts)) != AF_OK) goto out;
  t1 = now_ms();
  /* the device matrix to the host */
  {
    const long tot = s->K.total_slots * 64, nsl = s->K.n_slices;
    hc = (int *)malloc(sizeof(int) * (size_t)tot);
    hl = (int *)malloc(sizeof(int) * (size_t)nu);
    ho = (long *)malloc(sizeof(long) * (size_t)nsl);
    hv = (double *)malloc(sizeof(double) * (size_t)tot * 9);
    if (!hc || !hl || !ho || !hv) { rv = AF_ERR_MEM; goto out; }
    if (ugk_d2h(hc, s->K.d_cols, sizeof(int) * (size_t)tot, s->stream) ||
        ugk_d2h(hl, s->K.d_rowlen, sizeof(int) * (size_t)nu, s->stream) ||
        ugk_d2h(ho, s->K.d_soff, sizeof(long) * (size_t)nsl, s->stream) ||
        ugk_d2h(hv, s->K.d_vals, sizeof(double) * (size_t)tot * 9, s->stream)) {
      ufem_set_error("fill_K: matrix download: %s", ugk_error());
      rv = AF_ERR;
      goto out;
    }
  }
  /* ... and into the row slots, node rows split over up to 8 threads */
  {
    wb_ctx w;
    wb_job j;
    pthread_t th[8];
    long p = sysconf(_SC_NPROCESSORS_ONLN);
    int nt = nu < 4096 ? 1 : (int)(p < 1 ? 1 : (p > 8 ? 8 : p)), started = 0;
    const char *e = getenv("UFEM_HOST_THREADS");
    if (e && atoi(e) >= 1 && atoi(e) <= 8) nt = atoi(e);
    memset(&w, 0, sizeof(w));
    w.s = s;
    w.cols = hc;
    w.rowlen = hl;
    w.soff = ho;
    w.vals = hv;
    w.used = used;
    w.map = dof_map;
    w.known = known_dofs;
    w.K = K;
    j.w = &w;
    j.n = nu;
    j.next = 0;
    for (c = 1; c < nt; c++)
      if (pthread_create(&th[started], NULL, wb_worker, &j) == 0) started++;
    wb_worker(&j);
    for (c = 0; c < started; c++) pthread_join(th[c], NULL);
    rv = AF_OK;
    for (c = 0; c < 64; c++)
      if (w.err[c]) {
        rv = w.err[c];
        if (rv == AF_ERR_BIG) ufem_set_error("fill_K: row %ld has no free slot left", w.err_row[c]);
        else if (rv == AF_ERR_SIZ) ufem_set_error("fill_K: DOF map names row %ld of %ld", w.err_row[c], K->rows);
        else ufem_set_error("fill_K: out of host memory");
        break;
      }
  }
  ufem_gpu_mark_dirty(K);
  t2 = now_ms();
  g_fill.t[0] = (t1 - t0) - s->t_fill_ms;
  g_fill.t[1] = s->t_fill_ms;
  g_fill.t[2] = t2 - t1;
out:
  free(newid); free(used); free(en); free(x); free(y); free(z);
  free(hc); free(hl); free(ho); free(hv);
  return rv;
}
