/*
 * ufem_batch.c -- many independent linear-static problems (Monte-Carlo
 * samples of material fields on possibly different meshes) solved in one
 * persistent launch, plus the autotuner that picks its launch shape.
 *
 * The reference solves Monte-Carlo samples one after another through its
 * plugin API (monte_solution -> monte_linear_solver, fem/fem_mont.c:480-511,
 * :694-855), one mesh per library instance.  Here samples that share a mesh
 * (connectivity, coordinates, loads and supports) share one device pattern;
 * every sample owns its matrix values, Jacobi diagonal and Krylov vectors.
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "ufem_fem_math.h"
#include "ufem_internal.h"

typedef struct {
  ufem_sys *sys;          /* pattern, colours, mesh, supports, F of the class */
} batch_class;

struct ufem_batch {
  int n;                  /* samples */
  int n_cls;
  batch_class *cls;
  int *cls_of;            /* sample -> class */
  double **E_host, **nu_host;
  double **d_vals, **d_dinv, **d_vec;   /* per sample: K values, 1/diag, 6 vectors */
  float **d_vals32, **d_corr32;         /* per sample, mixed: FP32 copy and its correction */
  unsigned short **d_rem16;             /* per sample, mixed 2: bf16 remainders of the copy */
  int mixed;              /* 0 FP64 K; 1 the FP32 copies; 2 FP32 + bf16 remainders (48-bit) */
  int mixed_ready;        /* the mode the sample table holds copies for */
  ugk_bsample *h_samples, *d_samples;
  int *d_slots;           /* 2 ints per slot: sample, (team << 8) | rank */
  unsigned *d_ticket, *d_tctr;
  double *d_tpart;
  ugk_bresult *d_res, *h_res;
  void *stream, *ev0, *ev1;
  int assembled;
  int wg, grid;           /* launch shape (0: not tuned yet) */
  int team_big;           /* team size of the largest class */
  int *h_abort;           /* mapped host flag: the kernel stops when it is set */
  int *d_abort;
  double wait_ms;         /* a team wait gives up after this long */
  double timeout_ms;      /* the host sets h_abort when a solve takes longer (0: never) */
  int aborted;            /* the last solve was stopped */
};

static int gpu_fail(const char *what)
{
  ufem_set_error("%s: %s", what, ugk_error());
  return AF_ERR;
}

static int same_class(const ufem_mesh *a, const ufem_mesh *b)
{
  if (a->n_nodes != b->n_nodes || a->n_elems != b->n_elems || a->n_nl != b->n_nl) return 0;
  if (memcmp(a->enodes, b->enodes, sizeof(long) * (size_t)a->n_elems * 8)) return 0;
  if (memcmp(a->x, b->x, sizeof(double) * (size_t)a->n_nodes) ||
      memcmp(a->y, b->y, sizeof(double) * (size_t)a->n_nodes) ||
      memcmp(a->z, b->z, sizeof(double) * (size_t)a->n_nodes))
    return 0;
  if (a->n_nl > 0 &&
      (memcmp(a->nl_node, b->nl_node, sizeof(long) * (size_t)a->n_nl) ||
       memcmp(a->nl_type, b->nl_type, sizeof(long) * (size_t)a->n_nl) ||
       memcmp(a->nl_dir, b->nl_dir, sizeof(long) * (size_t)a->n_nl) ||
       memcmp(a->nl_val, b->nl_val, sizeof(double) * (size_t)a->n_nl)))
    return 0;
  return 1;
}

void ufem_batch_destroy(ufem_batch *b)
{
  int i;
  if (!b) return;
  if (b->stream) ugk_sync(b->stream);
  for (i = 0; i < b->n; i++) {
    if (b->d_vals) ugk_free(b->d_vals[i]);
    if (b->d_dinv) ugk_free(b->d_dinv[i]);
    if (b->d_vec) ugk_free(b->d_vec[i]);
    if (b->d_vals32) ugk_free(b->d_vals32[i]);
    if (b->d_corr32) ugk_free(b->d_corr32[i]);
    if (b->d_rem16) ugk_free(b->d_rem16[i]);
    if (b->E_host) free(b->E_host[i]);
    if (b->nu_host) free(b->nu_host[i]);
  }
  for (i = 0; i < b->n_cls; i++) ufem_sys_destroy(b->cls[i].sys);
  free(b->cls); free(b->cls_of); free(b->E_host); free(b->nu_host);
  free(b->d_vals); free(b->d_dinv); free(b->d_vec); free(b->h_samples); free(b->h_res);
  free(b->d_vals32); free(b->d_corr32); free(b->d_rem16);
  ugk_free(b->d_samples); ugk_free(b->d_slots); ugk_free(b->d_ticket); ugk_free(b->d_tctr);
  ugk_free(b->d_tpart); ugk_free(b->d_res);
  ugk_event_destroy(b->ev0);
  ugk_event_destroy(b->ev1);
  ugk_stream_destroy(b->stream);
  ugk_host_free(b->h_abort);
  free(b);
}

/* a team wait gives up after wait_ms (<= 0: 10 s); the host stops a solve
 * that has run timeout_ms (<= 0: never) by setting the kernel's stop flag */
void ufem_batch_set_limits(ufem_batch *b, double wait_ms, double timeout_ms)
{
  b->wait_ms = wait_ms > 0 ? wait_ms : 10000.0;
  b->timeout_ms = timeout_ms > 0 ? timeout_ms : 0.0;
}

/* set the stop flag of a running solve (from another host thread) */
void ufem_batch_abort(ufem_batch *b)
{
  if (b->h_abort) __atomic_store_n(b->h_abort, 1, __ATOMIC_SEQ_CST);
}

ufem_batch *ufem_batch_create(int n, const ufem_mesh *meshes)
{
  ufem_batch *b;
  int i, c;
  const ufem_mesh **rep;
  if (n < 1) { ufem_set_error("empty batch"); return NULL; }
  b = (ufem_batch *)calloc(1, sizeof(ufem_batch));
  rep = (const ufem_mesh **)calloc((size_t)n, sizeof(*rep));
  if (!b || !rep) { free(b); free(rep); return NULL; }
  b->n = n;
  b->cls = (batch_class *)calloc((size_t)n, sizeof(batch_class));
  b->cls_of = (int *)calloc((size_t)n, sizeof(int));
  b->E_host = (double **)calloc((size_t)n, sizeof(double *));
  b->nu_host = (double **)calloc((size_t)n, sizeof(double *));
  b->d_vals = (double **)calloc((size_t)n, sizeof(double *));
  b->d_dinv = (double **)calloc((size_t)n, sizeof(double *));
  b->d_vec = (double **)calloc((size_t)n, sizeof(double *));
  b->d_vals32 = (float **)calloc((size_t)n, sizeof(float *));
  b->d_corr32 = (float **)calloc((size_t)n, sizeof(float *));
  b->d_rem16 = (unsigned short **)calloc((size_t)n, sizeof(unsigned short *));
  b->h_samples = (ugk_bsample *)calloc((size_t)n, sizeof(ugk_bsample));
  b->h_res = (ugk_bresult *)calloc((size_t)n, sizeof(ugk_bresult));
  if (!b->cls || !b->cls_of || !b->E_host || !b->nu_host || !b->d_vals || !b->d_dinv || !b->d_vec ||
      !b->d_vals32 || !b->d_corr32 || !b->d_rem16 || !b->h_samples || !b->h_res)
    goto fail;
  for (i = 0; i < n; i++) {
    for (c = 0; c < b->n_cls; c++)
      if (same_class(rep[c], &meshes[i])) break;
    if (c == b->n_cls) {
      b->cls[c].sys = ufem_sys_create(&meshes[i]);
      if (!b->cls[c].sys) goto fail;
      /* the batch kernel solves K u = F with F zeroed at the supports: a
         support of non-zero value would be solved as if it were 0 */
      if (b->cls[c].sys->any_prescribed) {
        ufem_set_error("batch: sample %d has a support of non-zero value (solve it with ufem_sys_*)", i);
        b->n_cls++;
        goto fail;
      }
      /* loads and supports once per class (identical within a class) */
      if (ufem_sys_add_loads(b->cls[c].sys) != AF_OK) goto fail;
      rep[c] = &meshes[i];
      b->n_cls++;
    }
    b->cls_of[i] = c;
    b->E_host[i] = (double *)malloc(sizeof(double) * (size_t)meshes[i].n_elems);
    b->nu_host[i] = (double *)malloc(sizeof(double) * (size_t)meshes[i].n_elems);
    if (!b->E_host[i] || !b->nu_host[i]) goto fail;
    memcpy(b->E_host[i], meshes[i].E, sizeof(double) * (size_t)meshes[i].n_elems);
    memcpy(b->nu_host[i], meshes[i].nu, sizeof(double) * (size_t)meshes[i].n_elems);
  }
  b->stream = ugk_stream_create();
  b->ev0 = ugk_event_create();
  b->ev1 = ugk_event_create();
  b->d_samples = (ugk_bsample *)ugk_malloc(sizeof(ugk_bsample) * (size_t)n);
  b->d_slots = (int *)ugk_malloc(sizeof(int) * 2 * (size_t)n * 8);
  b->d_ticket = (unsigned *)ugk_malloc(64);
  b->d_tctr = (unsigned *)ugk_malloc(sizeof(unsigned) * (size_t)n);
  b->d_tpart = (double *)ugk_malloc(sizeof(double) * 64 * (size_t)n);
  b->d_res = (ugk_bresult *)ugk_malloc(sizeof(ugk_bresult) * (size_t)n);
  b->h_abort = (int *)ugk_host_alloc_mapped(64, (void **)&b->d_abort);
  b->wait_ms = 10000.0;
  b->team_big = 1;
  if (!b->stream || !b->ev0 || !b->ev1 || !b->d_samples || !b->d_slots || !b->d_ticket || !b->d_tctr ||
      !b->d_tpart || !b->d_res || !b->h_abort) {
    gpu_fail("batch allocation");
    goto fail;
  }
  for (i = 0; i < n; i++) {
    const ufem_sys *S = b->cls[b->cls_of[i]].sys;
    size_t nv = (size_t)S->K.total_slots * 64 * 9;
    b->d_vals[i] = (double *)ugk_malloc(sizeof(double) * nv);
    b->d_dinv[i] = (double *)ugk_malloc(sizeof(double) * (size_t)S->n);
    b->d_vec[i] = (double *)ugk_malloc(sizeof(double) * (size_t)S->n * 6);
    if (!b->d_vals[i] || !b->d_dinv[i] || !b->d_vec[i]) { gpu_fail("sample allocation"); goto fail; }
  }
  free(rep);
  return b;
fail:
  free(rep);
  ufem_batch_destroy(b);
  return NULL;
}

int ufem_batch_nclasses(const ufem_batch *b) { return b->n_cls; }

/* fem_fill_K(AF_NO) + fem_add_loads + fem_add_disps for every sample: the
 * element kernel scatters into the sample's own values on the class
 * pattern, the supports are applied on the device, and the class's F (loads
 * with supported DOFs zeroed) is shared. */
int ufem_batch_assemble(ufem_batch *b, int npts, double *ms)
{
  int i, c, first = 1;
  ugk_state st;
  if (ugk_event_record(b->ev0, b->stream)) return gpu_fail("event");
  for (i = 0; i < b->n; i++) {
    ufem_sys *S = b->cls[b->cls_of[i]].sys;
    ugk_mat K = S->K;
    K.d_vals = b->d_vals[i];
    K.d_vals32 = NULL;
    K.d_corr32 = NULL;
    K.d_rem16 = NULL;
    if (ugk_h2d(S->d_E, b->E_host[i], sizeof(double) * (size_t)S->n_elems, b->stream) ||
        ugk_h2d(S->d_nu, b->nu_host[i], sizeof(double) * (size_t)S->n_elems, b->stream) ||
        ugk_zero(K.d_vals, sizeof(double) * (size_t)K.total_slots * 64 * 9, b->stream))
      return gpu_fail("sample upload");
    for (c = 0; c < S->n_colors; c++) {
      long off = S->color_off[c], cnt = S->color_off[c + 1] - off;
      if (ugk_elem_stiff_scatter(&K, cnt, S->d_elist + off, S->d_x, S->d_y, S->d_z, S->d_enodes, S->d_E,
                                 S->d_nu, npts, S->w.st, b->stream))
        return gpu_fail("element stiffness");
    }
    if (S->n_fixed > 0 && ugk_apply_zero_bc(&K, S->n_fixed, S->d_fixed, b->stream)) return gpu_fail("supports");
    if (ugk_diag_inv3(&K, b->d_dinv[i], S->w.st, b->stream)) return gpu_fail("diagonal");
    /* the class F: supports zero F at fixed DOFs (host bookkeeping once) */
    if (!S->bc_applied) {
      long r;
      for (r = 0; r < S->n_react; r++) S->F_host[3 * S->react_node[r] + S->react_dof[r] - 1] = 0.0;
      if (ugk_h2d(S->d_F, S->F_host, sizeof(double) * (size_t)S->n, b->stream)) return gpu_fail("F");
      S->bc_applied = 1;
    }
    /* the E/nu upload of the next sample reuses the class buffers: keep order */
    if (ugk_sync(b->stream)) return gpu_fail("sync");
    (void)first;
  }
  if (ugk_event_record(b->ev1, b->stream)) return gpu_fail("event");
  if (ms) *ms = ugk_event_ms(b->ev0, b->ev1);
  for (c = 0; c < b->n_cls; c++) {
    if (ugk_d2h(&st, b->cls[c].sys->w.st, sizeof(st), b->stream)) return gpu_fail("state");
    if (st.err == 1) { ufem_set_error("element with non-positive Jacobian"); return AF_ERR_VAL; }
    if (st.err == 2) { ufem_set_error("zero diagonal"); return AF_ERR_ZER; }
    if (st.err) { ufem_set_error("assembly error %d", st.err); return AF_ERR; }
  }
  for (i = 0; i < b->n; i++) {
    const ufem_sys *S = b->cls[b->cls_of[i]].sys;
    ugk_bsample *q = &b->h_samples[i];
    double *v = b->d_vec[i];
    q->slen = S->K.d_slen;
    q->soff = S->K.d_soff;
    q->cols = S->K.d_cols;
    q->n_brows = S->K.n_brows;
    q->n_slices = S->K.n_slices;
    q->vals = b->d_vals[i];
    q->vals32 = NULL;   /* FP32 copies follow at the next solve in mixed mode */
    q->corr32 = NULL;
    q->rem16 = NULL;
    q->dinv = b->d_dinv[i];
    q->f = S->d_F;
    q->x = v;
    q->r = v + S->n;
    q->u = v + 2 * S->n;
    q->p = v + 3 * S->n;
    q->s = v + 4 * S->n;
    q->w = v + 5 * S->n;
  }
  if (ugk_h2d(b->d_samples, b->h_samples, sizeof(ugk_bsample) * (size_t)b->n, b->stream) || ugk_sync(b->stream))
    return gpu_fail("sample table");
  b->assembled = 1;
  b->mixed_ready = 0;
  return AF_OK;
}

/* the sample table's FP32 copies: made from the assembled values when the
 * batch is in mixed mode, left out otherwise */
static int mixed_table(ufem_batch *b)
{
  int i;
  const int want = b->mixed;
  if (want == b->mixed_ready) return AF_OK;
  for (i = 0; i < b->n; i++) {
    ugk_bsample *q = &b->h_samples[i];
    q->vals32 = NULL;
    q->corr32 = NULL;
    q->rem16 = NULL;
    if (want) {
      const ufem_sys *S = b->cls[b->cls_of[i]].sys;
      ugk_mat K = S->K;
      if (!b->d_vals32[i]) b->d_vals32[i] = (float *)ugk_malloc(sizeof(float) * (size_t)K.total_slots * 64 * 9);
      if (!b->d_corr32[i]) b->d_corr32[i] = (float *)ugk_malloc(sizeof(float) * 9 * (size_t)K.n_brows);
      if (want == 2 && !b->d_rem16[i])
        b->d_rem16[i] = (unsigned short *)ugk_malloc(sizeof(unsigned short) * (size_t)K.total_slots * 64 * 9);
      if (!b->d_vals32[i] || !b->d_corr32[i] || (want == 2 && !b->d_rem16[i])) return gpu_fail("fp32 copies");
      K.d_vals = b->d_vals[i];
      K.d_vals32 = b->d_vals32[i];
      K.d_corr32 = b->d_corr32[i];
      K.d_rem16 = want == 2 ? b->d_rem16[i] : NULL;
      if (ugk_vals_to_f32(&K, b->stream)) return gpu_fail("fp32 copies");
      q->vals32 = b->d_vals32[i];
      q->corr32 = b->d_corr32[i];
      q->rem16 = K.d_rem16;
    }
  }
  if (ugk_h2d(b->d_samples, b->h_samples, sizeof(ugk_bsample) * (size_t)b->n, b->stream) || ugk_sync(b->stream))
    return gpu_fail("sample table");
  b->mixed_ready = want;
  return AF_OK;
}

int ufem_batch_set_mixed(ufem_batch *b, int mode)
{
  if (!b) {
    ufem_set_error("batch mixed mode: no batch");
    return AF_ERR_VAL;
  }
  if (mode < 0 || mode > 2) {
    ufem_set_error("batch mixed mode %d: 0 (FP64), 1 (FP32 copy) or 2 (FP32 + bf16 remainder)", mode);
    return AF_ERR_VAL;
  }
  b->mixed = mode;
  return AF_OK;
}

/* Work of a Jacobi-PCG solve on these meshes grows like n^(4/3) (rows times
 * an iteration count that grows with the mesh length).  The largest class
 * gets teams of `team_big` workgroups, smaller classes proportionally fewer
 * (powers of two); slots are dealt largest work-per-workgroup first
 * (longest-processing-time order), a team's slots adjacent. */
static int build_slots(ufem_batch *b, int team_big, int grid, int budget, int *nslots_out)
{
  int i, j, ns = 0, *ord = NULL, *slots = NULL, *team = NULL;
  double *key = NULL;
  long nmax = 1;
  for (i = 0; i < b->n; i++) if (b->h_samples[i].n_brows > nmax) nmax = b->h_samples[i].n_brows;
  ord = (int *)malloc(sizeof(int) * (size_t)b->n);
  team = (int *)malloc(sizeof(int) * (size_t)b->n);
  key = (double *)malloc(sizeof(double) * (size_t)b->n);
  slots = (int *)malloc(sizeof(int) * 2 * (size_t)b->n * 8);
  if (!ord || !team || !key || !slots) { free(ord); free(team); free(key); free(slots); return AF_ERR_MEM; }
  if (team_big > 8) team_big = 8;
  if (team_big > grid) team_big = grid;
  for (i = 0; i < b->n; i++) {
    double f = (double)b->h_samples[i].n_brows / (double)nmax;
    int t = 1;
    while (t * 2 <= team_big && (double)(t * 2) <= team_big * f * 1.5) t *= 2;
    team[i] = t;
    key[i] = pow((double)b->h_samples[i].n_brows, 4.0 / 3.0) / t;
    ord[i] = i;
    b->h_samples[i].maxit = budget > 0 ? (long)ceil(budget * pow(f, 1.0 / 3.0)) : 0;
  }
  for (i = 1; i < b->n; i++) {   /* stable insertion sort, descending key */
    int v = ord[i];
    for (j = i - 1; j >= 0 && key[ord[j]] < key[v]; j--) ord[j + 1] = ord[j];
    ord[j + 1] = v;
  }
  for (i = 0; i < b->n; i++) {
    int id = ord[i], r;
    for (r = 0; r < team[id]; r++) {
      slots[2 * ns] = id;
      slots[2 * ns + 1] = (team[id] << 8) | r;
      ns++;
    }
  }
  i = ugk_h2d(b->d_slots, slots, sizeof(int) * 2 * (size_t)ns, b->stream) ||
      ugk_h2d(b->d_samples, b->h_samples, sizeof(ugk_bsample) * (size_t)b->n, b->stream);
  free(ord); free(team); free(key); free(slots);
  if (i) return gpu_fail("slots");
  *nslots_out = ns;
  return AF_OK;
}

static double wall_ms(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return 1e3 * (double)ts.tv_sec + 1e-6 * (double)ts.tv_nsec;
}

static int run_batch(ufem_batch *b, double eps, long maxit, int wg, int grid, int team_big, int budget,
                     double *ms)
{
  int ns = 0, i;
  const long khz = ugk_wall_clock_khz();
  const long wait_ticks = khz > 0 ? (long)ceil(b->wait_ms * (double)khz) : 0;   /* at least 1 tick */
  if (build_slots(b, team_big, grid, budget, &ns) != AF_OK) return AF_ERR;
  __atomic_store_n(b->h_abort, 0, __ATOMIC_SEQ_CST);
  b->aborted = 0;
  memset(b->h_res, 0, sizeof(ugk_bresult) * (size_t)b->n);
  if (ugk_zero(b->d_ticket, 64, b->stream) || ugk_zero(b->d_tctr, sizeof(unsigned) * (size_t)b->n, b->stream) ||
      ugk_h2d(b->d_res, b->h_res, sizeof(ugk_bresult) * (size_t)b->n, b->stream) ||
      ugk_event_record(b->ev0, b->stream) ||
      ugk_batch_pcg(b->d_samples, b->d_slots, ns, b->d_ticket, b->d_tctr, b->d_tpart, b->d_res, eps, maxit, wg,
                    grid, b->d_abort, wait_ticks, b->stream) ||
      ugk_event_record(b->ev1, b->stream))
    return gpu_fail("batched solve");
  if (b->timeout_ms > 0) {            /* the host gives up a bounded wait */
    const double t0 = wall_ms();
    int d;
    while ((d = ugk_event_done(b->ev1)) == 0) {
      if (wall_ms() - t0 > b->timeout_ms) {
        __atomic_store_n(b->h_abort, 1, __ATOMIC_SEQ_CST);
        b->aborted = 1;
        break;
      }
      struct timespec nap = {0, 200000};
      nanosleep(&nap, NULL);
    }
    if (d < 0) return gpu_fail("batched solve");
  }
  if (ugk_sync(b->stream)) return gpu_fail("batched solve");
  *ms = ugk_event_ms(b->ev0, b->ev1);
  if (ugk_d2h(b->h_res, b->d_res, sizeof(ugk_bresult) * (size_t)b->n, b->stream)) return gpu_fail("results");
  for (i = 0; i < b->n; i++)
    if (b->h_res[i].status < 1 || b->h_res[i].status > 3) b->aborted = 1;
  if (b->aborted) {
    ufem_set_error("batch: stopped before every sample was solved (%s)",
                   __atomic_load_n(b->h_abort, __ATOMIC_SEQ_CST) ? "the host gave up waiting" : "a team wait gave up");
    return AF_ERR;
  }
  return AF_OK;
}

/* Autotuner: times every candidate launch shape -- workgroup size (256, 512,
 * 1024), workgroups per CU (1, 2 where they fit) and the team size of the
 * largest class (1, 2, 4, 8) -- on a budget of iterations per sample scaled
 * like the iteration counts (budget x (n / n_max)^(1/3)), and keeps the
 * fastest.  Results are cached in a text file keyed by device and batch
 * shape, so later runs skip the search. */
int ufem_batch_autotune(ufem_batch *b, const char *cache, int budget_iters, int *wg_out, int *grid_out)
{
  static const int wgs[3] = {256, 512, 1024};
  static const int per[2] = {1, 2};
  static const int teams[4] = {1, 2, 4, 8};
  char dev[128] = "", key[256];
  int ncu = 0, best_wg = 0, best_grid = 0, best_team = 1, a, k, t, i;
  long hbm = 0, rows = 0;
  double best = 1e30, ms;
  if (!b->assembled) { ufem_set_error("assemble first"); return AF_ERR; }
  if (mixed_table(b) != AF_OK) return AF_ERR;
  if (ugk_device_info(&ncu, &hbm, dev, sizeof(dev))) return gpu_fail("device");
  for (i = 0; i < b->n; i++) rows += b->h_samples[i].n_brows;
  snprintf(key, sizeof(key), "%s|n=%d|rows=%ld|cls=%d|budget=%d%s", dev, b->n, rows, b->n_cls, budget_iters,
           b->mixed == 2 ? "|mixed48" : (b->mixed ? "|mixed" : ""));
  if (cache) {
    FILE *f = fopen(cache, "r");
    char line[512];
    if (f) {
      while (fgets(line, sizeof(line), f)) {
        char *tab = strrchr(line, '\t');
        if (tab && (size_t)(tab - line) == strlen(key) && !strncmp(line, key, strlen(key)) &&
            sscanf(tab + 1, "%d %d %d", &best_wg, &best_grid, &best_team) == 3) {
          fclose(f);
          b->wg = best_wg;
          b->grid = best_grid;
          b->team_big = best_team;
          if (wg_out) *wg_out = best_wg;
          if (grid_out) *grid_out = best_grid;
          return AF_OK;
        }
      }
      fclose(f);
    }
  }
  for (a = 0; a < 3; a++) {
    int occ = ugk_batch_occupancy(wgs[a]);
    for (k = 0; k < 2; k++) {
      int grid = per[k] * ncu;
      if (per[k] > occ) continue;
      for (t = 0; t < 4; t++) {
        if (run_batch(b, 0.0, budget_iters, wgs[a], grid, teams[t], budget_iters, &ms) != AF_OK) return AF_ERR;
        if (ms < best) { best = ms; best_wg = wgs[a]; best_grid = grid; best_team = teams[t]; }
      }
    }
  }
  for (i = 0; i < b->n; i++) b->h_samples[i].maxit = 0;
  if (best_wg == 0) { ufem_set_error("no launch shape fits"); return AF_ERR; }
  b->wg = best_wg;
  b->grid = best_grid;
  b->team_big = best_team;
  if (cache) {
    FILE *f = fopen(cache, "a");
    if (f) {
      fprintf(f, "%s\t%d %d %d\n", key, best_wg, best_grid, best_team);
      fclose(f);
    }
  }
  if (wg_out) *wg_out = best_wg;
  if (grid_out) *grid_out = best_grid;
  return AF_OK;
}

int ufem_batch_solve(ufem_batch *b, double eps, long maxit, double *total_ms, long *iters, double *relres,
                     int *status, double *sample_ms)
{
  double ms = 0.0, khz;
  int i, wg = b->wg ? b->wg : 512, grid = b->grid;
  long mx = maxit;
  if (!b->assembled) { ufem_set_error("assemble first"); return AF_ERR; }
  if (mixed_table(b) != AF_OK) return AF_ERR;
  if (mx <= 0) {
    for (i = 0; i < b->n; i++) if (3 * b->h_samples[i].n_brows > mx) mx = 3 * b->h_samples[i].n_brows;
  }
  if (grid <= 0) {
    int ncu = 0;
    long hbm;
    char dev[64];
    if (ugk_device_info(&ncu, &hbm, dev, sizeof(dev))) return gpu_fail("device");
    grid = ncu * (ugk_batch_occupancy(wg) >= 2 ? 2 : 1);
  }
  {
    int rv = run_batch(b, eps, mx, wg, grid, b->team_big, 0, &ms);
    if (rv != AF_OK && !b->aborted) return rv;   /* a stopped solve still reports its samples */
  }
  if (total_ms) *total_ms = ms;
  khz = (double)ugk_wall_clock_khz();
  for (i = 0; i < b->n; i++) {
    const ugk_bresult *R = &b->h_res[i];
    if (iters) iters[i] = R->iters;
    if (relres) relres[i] = R->relres;
    if (status) status[i] = R->status;
    if (sample_ms) sample_ms[i] = khz > 0 ? (double)(R->t_end - R->t_start) / khz : -1.0;
  }
  for (i = 0; i < b->n; i++) if (b->h_res[i].status != 1) return AF_ERR;
  return AF_OK;
}

int ufem_batch_teams(const ufem_batch *b, int *team)
{
  int i;
  for (i = 0; i < b->n; i++) team[i] = b->h_res[i].team;
  return AF_OK;
}

int ufem_batch_get_u(ufem_batch *b, int i, double *u)
{
  const ufem_sys *S;
  if (i < 0 || i >= b->n) return AF_ERR_VAL;
  S = b->cls[b->cls_of[i]].sys;
  if (ugk_d2h(u, b->h_samples[i].x, sizeof(double) * (size_t)S->n, b->stream)) return gpu_fail("u");
  return AF_OK;
}

long ufem_batch_ndof(const ufem_batch *b, int i)
{
  if (i < 0 || i >= b->n) return -1;
  return b->cls[b->cls_of[i]].sys->n;
}

void ufem_batch_shape(const ufem_batch *b, int *wg, int *grid, int *team_big)
{
  if (wg) *wg = b->wg;
  if (grid) *grid = b->grid;
  if (team_big) *team_big = b->team_big;
}

void ufem_batch_set_shape(ufem_batch *b, int wg, int grid, int team_big)
{
  b->wg = wg;
  b->grid = grid;
  b->team_big = team_big < 1 ? 1 : (team_big > 8 ? 8 : team_big);
}
