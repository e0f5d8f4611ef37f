/*
 * ufem_solve.c -- host control of the device solvers: Jacobi PCG (streamed
 * graph-replayed kernels or one persistent cooperative launch), Jacobi
 * BiCGStab, and FP32-inner / FP64-refinement mixed precision.
 *
 * The iteration itself lives on the device (ugk_kernels.hip); the host only
 * enqueues work and, for the streamed form, polls the device's `done` flag
 * once per chunk of iterations while the next chunk is already queued.
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "ufem_fem_math.h"
#include "ufem_internal.h"

static int gpu_fail(const char *what)
{
  ufem_set_error("%s: %s", what, ugk_error());
  return AF_ERR;
}

void ufem_work_free(ufem_work *w)
{
  ugk_free(w->r); ugk_free(w->z); ugk_free(w->p[0]); ugk_free(w->p[1]); ugk_free(w->q);
  ugk_free(w->dinv); ugk_free(w->rhat); ugk_free(w->phat); ugk_free(w->v); ugk_free(w->s);
  ugk_free(w->shat); ugk_free(w->t); ugk_free(w->r32); ugk_free(w->z32); ugk_free(w->p32[0]);
  ugk_free(w->p32[1]); ugk_free(w->q32); ugk_free(w->x32); ugk_free(w->dinv32);
  ugk_free(w->fP); ugk_free(w->fQ);
  ufem_ssor_free(w->ssor);
  ufem_coarse_free(w->cc);
  ugk_free(w->part); ugk_free(w->st);
  ugk_host_free(w->st_host);
  memset(w, 0, sizeof(*w));
}

int ufem_work_alloc(ufem_work *w, long n, int with_bicgs, int with_mixed)
{
  size_t b = sizeof(double) * (size_t)n, f = sizeof(float) * (size_t)n;
  memset(w, 0, sizeof(*w));
  w->n = n;
  w->r = (double *)ugk_malloc(b);
  w->z = (double *)ugk_malloc(b);
  w->p[0] = (double *)ugk_malloc(b);
  w->p[1] = (double *)ugk_malloc(b);
  w->q = (double *)ugk_malloc(b);
  w->dinv = (double *)ugk_malloc(b);
  w->part = (double *)ugk_malloc(sizeof(double) * (size_t)UGK_PART_DOUBLES);
  w->st = (ugk_state *)ugk_malloc(sizeof(ugk_state));
  w->st_host = (ugk_state *)ugk_host_alloc(sizeof(ugk_state));
  if (!w->r || !w->z || !w->p[0] || !w->p[1] || !w->q || !w->dinv || !w->part || !w->st || !w->st_host)
    goto fail;
  if (with_bicgs) {
    w->rhat = (double *)ugk_malloc(b);
    w->phat = (double *)ugk_malloc(b);
    w->v = (double *)ugk_malloc(b);
    w->s = (double *)ugk_malloc(b);
    w->shat = (double *)ugk_malloc(b);
    w->t = (double *)ugk_malloc(b);
    if (!w->rhat || !w->phat || !w->v || !w->s || !w->shat || !w->t) goto fail;
  }
  if (with_mixed) {
    w->r32 = (float *)ugk_malloc(f);
    w->z32 = (float *)ugk_malloc(f);
    w->p32[0] = (float *)ugk_malloc(f);
    w->p32[1] = (float *)ugk_malloc(f);
    w->q32 = (float *)ugk_malloc(f);
    w->x32 = (float *)ugk_malloc(f);
    w->dinv32 = (float *)ugk_malloc(f);
    if (!w->r32 || !w->z32 || !w->p32[0] || !w->p32[1] || !w->q32 || !w->x32 || !w->dinv32) goto fail;
  }
  return AF_OK;
fail:
  gpu_fail("workspace allocation");
  ufem_work_free(w);
  return AF_ERR_MEM;
}

/* host copy of the scalar state, written before a solve */
static int put_state(ufem_work *w, void *stream, double tol, int rule, long maxit)
{
  ugk_state *h = w->st_host;
  memset(h, 0, sizeof(*h));
  h->tol = tol;
  h->rule = rule;
  h->maxit = maxit;
  if (ugk_h2d(w->st, h, sizeof(*h), stream)) return gpu_fail("state upload");
  return AF_OK;
}

static int get_state(ufem_work *w, void *stream)
{
  if (ugk_d2h(w->st_host, w->st, sizeof(ugk_state), stream)) return gpu_fail("state read");
  return AF_OK;
}

/* Enqueue chunks of iterations until the device reports done.  The state of
 * chunk k is copied to pinned memory behind an event while chunk k+1 is
 * queued, so the GPU never waits on the host. */
typedef int (*iter_fn)(void *ctx, long it, void *stream);

static int run_chunks(ufem_work *w, void *stream, long maxit, int chunk, iter_fn fn, void *ctx,
                      int use_graph)
{
  void *graph = NULL, *ev = NULL;
  long it = 0;
  int rv = AF_OK, first = 1;
  if (chunk < 2) chunk = 2;
  chunk += chunk & 1;                    /* even: ping-pong parity repeats */
  {                                      /* UFEM_GRAPHS=0: direct launches (kernel traces) */
    const char *e = getenv("UFEM_GRAPHS");
    if (e && e[0] == '0') use_graph = 0;
  }
  ev = ugk_event_create();
  if (!ev) return gpu_fail("event");
  if (w->hook && w->hook(w->hook_ctx, 0, stream)) { rv = gpu_fail("chunk hook"); goto out; }
  if (use_graph) {
    long k;
    if (ugk_graph_begin(stream)) { rv = gpu_fail("graph capture"); goto out; }
    for (k = 0; k < chunk; k++)
      if (fn(ctx, k, stream)) { (void)ugk_graph_end(stream); rv = gpu_fail("iteration"); goto out; }
    graph = ugk_graph_end(stream);
    if (!graph) use_graph = 0;           /* fall back to direct launches */
  }
  for (;;) {
    long k;
    if (use_graph) {
      if (ugk_graph_launch(graph, stream)) { rv = gpu_fail("graph launch"); goto out; }
    } else {
      for (k = 0; k < chunk; k++)
        if (fn(ctx, it + k, stream)) { rv = gpu_fail("iteration"); goto out; }
    }
    it += chunk;
    if (w->hook && w->hook(w->hook_ctx, it, stream)) { rv = gpu_fail("chunk hook"); goto out; }
    if (!first) {
      if (ugk_event_sync(ev)) { rv = gpu_fail("event sync"); goto out; }
      if (w->st_host->done) break;
    }
    first = 0;
    if (ugk_d2h_async(w->st_host, w->st, sizeof(ugk_state), stream) || ugk_event_record(ev, stream)) {
      rv = gpu_fail("state poll");
      goto out;
    }
    if (it > maxit + 2L * chunk) {       /* the device stops itself at maxit */
      if (ugk_event_sync(ev)) { rv = gpu_fail("event sync"); goto out; }
      break;
    }
  }
  if (ugk_sync(stream)) rv = gpu_fail("sync");
out:
  ugk_graph_destroy(graph);
  ugk_event_destroy(ev);
  return rv;
}

typedef struct {
  const ugk_mat *K;
  long n;
  double *x;
  ufem_work *w;
  int variant;
} pcg_ctx;

static int pcg_step(void *c, long it, void *stream)
{
  pcg_ctx *p = (pcg_ctx *)c;
  ufem_work *w = p->w;
  (void)it;
  if (p->variant == 2)   /* single reduction: u = z, p = p[0], s = p[1], w = q */
    return ugk_cg1_iter(p->K, p->n, p->x, w->r, w->z, w->p[0], w->p[1], w->q, w->dinv, w->st, w->part, stream);
  return ugk_pcg_iter(p->K, p->n, p->x, w->r, w->z, w->p, w->q, w->dinv, w->st, w->part, p->variant,
                      stream);
}

typedef struct {
  const ugk_mat *K;
  long n;
  const double *b;
  double *x;
  ufem_work *w;
} mx_ctx;

static int mx_step(void *c, long it, void *stream)
{
  mx_ctx *m = (mx_ctx *)c;
  ufem_work *w = m->w;
  (void)it;
  return ugk_mx_iter(m->K, m->n, m->b, m->x, w->r, w->r32, w->z32, w->p32[0], w->q32, w->x32, w->dinv32, w->st,
                     w->part, stream);
}

static int bicgs_step(void *c, long it, void *stream)
{
  pcg_ctx *p = (pcg_ctx *)c;
  ufem_work *w = p->w;
  (void)it;
  return ugk_bicgs_iter(p->K, p->n, p->x, w->r, w->rhat, w->p[0], w->phat, w->v, w->s, w->shat, w->t,
                        w->dinv, w->st, w->part, stream);
}

static int status_of(int done)
{
  return done == 1 ? AF_OK : AF_ERR;
}

/* FP64 true residual ||b - K x|| / ||b|| into info */
static int true_residual(const ugk_mat *K, long n, const double *b, double *x, ufem_work *w, void *stream,
                         double bb, double *out)
{
  if (ugk_residual(K, n, b, x, w->r, w->st, w->part, stream) || get_state(w, stream))
    return gpu_fail("residual");
  *out = bb > 0 ? sqrt(w->st_host->rr / bb) : 0.0;
  return AF_OK;
}

static int prepare(const ugk_mat *K, ufem_work *w, void *stream, double tol, int rule, long maxit)
{
  if (put_state(w, stream, tol, rule, maxit)) return AF_ERR;
  if (ugk_diag_inv(K, w->dinv, w->st, stream)) return gpu_fail("diagonal");
  if (rule == 0 && ugk_mat_norm(K, w->st, w->part, stream)) return gpu_fail("matrix norm");
  if (get_state(w, stream)) return AF_ERR;
  if (w->st_host->err == 2) { ufem_set_error("zero diagonal"); return AF_ERR_ZER; }
  return AF_OK;
}

static int solve_pcg(const ugk_mat *K, long n, const double *b, double *x, ufem_work *w, void *stream,
                     const ufem_solve_opts *o, long maxit, ufem_solve_info *info)
{
  int rv;
  pcg_ctx c = {K, n, x, w,
               o->exec == UFEM_EXEC_FUSED ? 0 : (o->exec == UFEM_EXEC_STREAMED ? 2 : 1)};
  rv = prepare(K, w, stream, o->eps, o->rule, maxit);
  if (rv != AF_OK) return rv;
  if (c.variant == 2) {
    if (ugk_cg1_start(K, n, b, x, w->r, w->z, w->p[0], w->p[1], w->dinv, w->part, stream))
      return gpu_fail("start");
  } else if (ugk_pcg_start(K, n, b, x, w->r, w->z, w->p[0], w->p[1], w->dinv, w->st, w->part, stream))
    return gpu_fail("start");
  if (o->exec == UFEM_EXEC_PERSISTENT) {
    if (ugk_pcg_persistent(K, n, x, w->r, w->z, w->p[0], w->p[1], w->q, w->dinv, w->st, w->part, stream))
      return gpu_fail("persistent PCG");
  } else {
    rv = run_chunks(w, stream, maxit, o->chunk > 0 ? o->chunk : 32, pcg_step, &c, 1);
    if (rv != AF_OK) return rv;
  }
  if (get_state(w, stream)) return AF_ERR;
  if (w->st_host->err == 5) {
    ufem_set_error("persistent PCG: a block did not reach the grid barrier (gave up)");
    return AF_ERR;
  }
  if (w->st_host->bb <= 0.0) {   /* zero load: the reference returns x = 0 */
    if (ugk_zero(x, sizeof(double) * (size_t)n, stream)) return gpu_fail("zero x");
    info->status = AF_OK;
    return AF_OK;
  }
  info->iters = w->st_host->iter;
  info->relres = sqrt(w->st_host->rr / w->st_host->bb);
  info->status = status_of(w->st_host->done);
  return info->status;
}

/* BiCGStab stops on its recursive residual, which drifts from b - A x as
 * the iterations go on.  A stop is therefore checked on the FP64 true
 * residual; if that misses the tolerance, BiCGStab restarts from the
 * current x (r = b - A x, new shadow residual) with the iterations left, at
 * most four times. */
static int solve_bicgs(const ugk_mat *K, long n, const double *b, double *x, ufem_work *w, void *stream,
                       const ufem_solve_opts *o, long maxit, ufem_solve_info *info)
{
  int rv, round;
  long total = 0;
  double rr_true = 0.0;
  pcg_ctx c = {K, n, x, w, 0};
  ugk_state *h = w->st_host;
  for (round = 0;; round++) {
    /* BiCGStab tests ||r|| <= eps ||b||; the reference rule is mapped on the
       same relative test with eps (its own test is not usable, see header) */
    rv = prepare(K, w, stream, o->eps, 1, maxit - total);
    if (rv != AF_OK) return rv;
    if (ugk_bicgs_start(K, n, b, x, w->r, w->rhat, w->p[0], w->v, w->st, w->part, stream))
      return gpu_fail("start");
    if (get_state(w, stream)) return AF_ERR;
    if (!h->done) {
      rv = run_chunks(w, stream, maxit - total, o->chunk > 0 ? o->chunk : 16, bicgs_step, &c, 1);
      if (rv != AF_OK) return rv;
      if (get_state(w, stream)) return AF_ERR;
    }
    total += h->iter;
    if (h->bb <= 0.0) {
      if (ugk_zero(x, sizeof(double) * (size_t)n, stream)) return gpu_fail("zero x");
      info->status = AF_OK;
      return AF_OK;
    }
    info->relres = sqrt(h->rr / h->bb);
    info->status = status_of(h->done);
    if (h->done != 1) break;
    {
      const double bb = h->bb;
      if (ugk_residual(K, n, b, x, w->r, w->st, w->part, stream) || get_state(w, stream))
        return gpu_fail("residual");
      rr_true = h->rr;
      if (sqrt(rr_true) <= o->eps * sqrt(bb)) break;
      info->status = AF_ERR;            /* a stop the true residual does not confirm */
      if (round >= 4 || total >= maxit) break;
    }
  }
  info->iters = total;
  info->outer = round;                  /* restarts */
  return info->status;
}

/* Mixed precision: K stored once more in FP32; CG runs on the FP32 matrix
 * and FP32 vectors (dot products summed in FP64), and whenever its recursive
 * residual has dropped by inner_eps (default 0.1) below the last true
 * residual -- or inner_maxit (default 500) iterations have passed -- the
 * FP32 solution increment is added to the FP64 solution and r = b - K u is
 * recomputed with the FP64 matrix.  By default the FP32 CG keeps its search
 * direction across such a refresh (reliable update); with inner_restart it
 * starts afresh on the correction equation K d = r, which is classical
 * iterative refinement with an inner CG to inner_eps (inner_maxit then
 * defaults to no cap).  The solve stops only on an FP64 residual,
 * ||b - K u|| <= eps ||b||, or after five refreshes in a row without
 * progress (AF_ERR). */
static int solve_mixed(ugk_mat *Km, long n, const double *b, double *x, ufem_work *w, void *stream,
                       const ufem_solve_opts *o, long maxit, ufem_solve_info *info)
{
  const ugk_mat *K = Km;
  ugk_state *h = w->st_host;
  mx_ctx c = {K, n, b, x, w};
  int rv;
  if ((rv = ufem_mat_f32(Km, 0, stream)) != AF_OK) return rv;
  rv = prepare(K, w, stream, o->eps, 1, maxit);
  if (rv != AF_OK) return rv;
  if (ugk_to_f32(n, w->dinv, w->dinv32, stream)) return gpu_fail("fp32 diagonal");
  memset(h, 0, sizeof(*h));
  h->tol = o->eps;
  h->rule = 1;
  h->maxit = maxit;
  h->delta = o->inner_eps > 0 ? o->inner_eps : 0.1;
  h->ru_req = -1;
  h->ru_at = 0;
  /* classical refinement runs each inner CG to inner_eps, uncapped by default */
  h->max_gap = o->inner_maxit > 0 ? o->inner_maxit : (o->inner_restart ? maxit : 500);
  h->restart = o->inner_restart ? 1 : 0;
  if (ugk_h2d(w->st, h, sizeof(*h), stream)) return gpu_fail("state");
  if (ugk_mx_start(K, n, b, x, w->r, w->r32, w->z32, w->p32[0], w->x32, w->dinv32, w->part, stream))
    return gpu_fail("mixed start");
  rv = run_chunks(w, stream, maxit, o->chunk > 0 ? o->chunk : 32, mx_step, &c, 1);
  if (rv != AF_OK) return rv;
  if (get_state(w, stream)) return AF_ERR;
  if (h->bb <= 0.0) {
    if (ugk_zero(x, sizeof(double) * (size_t)n, stream)) return gpu_fail("zero x");
    info->status = AF_OK;
    return AF_OK;
  }
  info->iters = h->iter;
  info->outer = h->n_ru;
  info->relres = sqrt(h->rr_true / h->bb);
  info->status = status_of(h->done);
  return info->status;
}

/* Flexible CG with the learned preconditioner (o->net bound to this system):
 * per iteration the head kernel tests convergence and writes the network
 * input, the network runs on the same stream (skipped through st->halt once
 * done or after the Jacobi fallback), then z = dinv r + s mask N, the
 * Polak-Ribiere direction update, q = K p and the x, r update.  The whole
 * iteration is captured in the chunk graph like the PCG kernels. */
typedef struct {
  const ugk_mat *K;
  long n;
  const double *b;
  double *x;
  ufem_work *w;
  ufem_net *net;             /* NULL: the coarse correction alone          */
  const ugk_coarse *cc;      /* NULL: the network alone                    */
  int m;
  int mixed;
  int fused;                 /* m = 1, coarse-first network: last layer in the combine */
  int f64;                   /* fused, with FP64 inference */
  ugk_conv_desc last;
} fcg_ctx;

static int fcg_step(void *c, long it, void *stream)
{
  fcg_ctx *f = (fcg_ctx *)c;
  ufem_work *w = f->w;
  (void)it;
  if (f->mixed && ugk_fcg_refresh(f->K, f->b, f->x, w->r, w->dinv, w->st, w->part, stream)) return -1;
  if (ugk_fcg_head(f->K, f->n, w->r, w->dinv, f->net ? ufem_net_input(f->net) : NULL, w->st, w->part, stream) ||
      (f->net && ufem_net_infer(f->net, &w->st->halt, stream)) ||
      (f->cc && ugk_cc_apply(f->cc, f->n, w->r, f->cc->d_zc, w->st, stream)) ||
      ugk_fcg_combine(f->n, w->r, w->dinv, f->net ? ufem_net_mask(f->net) : NULL,
                      f->net ? ufem_net_output(f->net) : NULL, w->fP, w->fQ, f->m, f->cc ? f->cc->d_zc : NULL,
                      w->z, w->st, w->part, stream) ||
      ugk_fcg_p(f->n, w->r, w->dinv, w->z, w->fP, f->m, w->st, w->part, stream) ||
      ugk_fcg_q(f->K, f->n, w->fP, w->fQ, f->m, f->mixed, w->st, w->part, stream))
    return -1;
  return ugk_fcg_update(f->K, f->n, f->x, w->r, w->fP, w->fQ, f->m, w->dinv, w->st, w->part, stream);
}

/* one kept direction: [refresh], network, [coarse term], combine, w = K z,
 * update (which also writes the next network input); p and q in ring slot
 * 0, w in the workspace's q */
static int fcg1_step(void *c, long it, void *stream)
{
  fcg_ctx *f = (fcg_ctx *)c;
  ufem_work *w = f->w;
  unsigned short *x0 = f->net && !f->f64 ? ufem_net_input(f->net) : NULL;
  double *v64 = f->f64 ? ufem_net_input64(f->net) : NULL;
  const float *mk = f->net ? ufem_net_mask(f->net) : NULL;
  (void)it;
  if (f->mixed &&
      ugk_f1_refresh(f->K, f->n, f->b, f->x, w->r, w->dinv, w->fP, w->fQ, x0, v64, mk, w->st, w->part, stream))
    return -1;
  if (f->fused) {                       /* the network's last layer inside the combine */
    if ((f->f64 ? ufem_net_infer64_head(f->net, &w->st->halt, stream)
                : ufem_net_infer_head(f->net, &w->st->halt, stream)) ||
        (f->cc && ugk_cc_apply(f->cc, f->n, w->r, f->cc->d_zc, w->st, stream)) ||
        ugk_f1_combine_c0t(f->n, w->r, w->dinv, ufem_net_mask(f->net), &f->last, f->cc ? f->cc->d_zc : NULL, w->fQ,
                           w->z, w->st, w->part, stream))
      return -1;
  } else if ((f->net && ufem_net_infer(f->net, &w->st->halt, stream)) ||
             (f->cc && ugk_cc_apply(f->cc, f->n, w->r, f->cc->d_zc, w->st, stream)) ||
             ugk_f1_combine(f->n, w->r, w->dinv, f->net ? ufem_net_mask(f->net) : NULL,
                            f->net ? ufem_net_output(f->net) : NULL, f->cc ? f->cc->d_zc : NULL, w->fQ, w->z, w->st,
                            w->part, stream))
    return -1;
  if (ugk_f1_w(f->K, w->z, w->q, f->mixed, w->st, w->part, stream)) return -1;
  return ugk_f1_update(f->K, f->n, f->x, w->r, w->fP, w->fQ, w->z, w->q, w->dinv, x0, v64, mk, w->st, w->part,
                       stream);
}

static int solve_learned(ugk_mat *Km, long n, const double *b, double *x, ufem_work *w, void *stream,
                         const ufem_solve_opts *o, long maxit, int mixed, ufem_solve_info *info)
{
  const ugk_mat *K = Km;
  ufem_net *N = (ufem_net *)o->net;
  const int m = o->fcg_m > 0 ? o->fcg_m : 1;   /* the network is a fixed SPD map */
  const ugk_coarse *cc = o->coarse ? w->cc : NULL;
  fcg_ctx c = {K, n, b, x, w, N, cc, m, mixed, 0, 0, {0}};
  ugk_state *h = w->st_host;
  int rv;
  if (o->coarse && !w->cc) {
    ufem_set_error("learned solver: coarse correction asked for without a coarse grid (ufem_sys_set_coarse)");
    return AF_ERR_VAL;
  }
  if (N ? ufem_net_bound_n(N) != n : !cc) {
    ufem_set_error("learned solver: no network bound to this system");
    return AF_ERR_VAL;
  }
  if (m > 8) { ufem_set_error("learned solver: at most 8 kept directions"); return AF_ERR_VAL; }
  if (w->fm < m) {
    ugk_free(w->fP);
    ugk_free(w->fQ);
    w->fP = (double *)ugk_malloc(sizeof(double) * (size_t)n * (size_t)m);
    w->fQ = (double *)ugk_malloc(sizeof(double) * (size_t)n * (size_t)m);
    w->fm = (w->fP && w->fQ) ? m : 0;
    if (!w->fm) return gpu_fail("direction rings");
  }
  if (mixed) {
    if ((rv = ufem_mat_f32(Km, o->f32_rem, stream)) != AF_OK) return rv;
  }
  rv = prepare(K, w, stream, o->eps, o->rule, maxit);
  if (rv != AF_OK) return rv;
  if (mixed) {
    h->delta = o->inner_eps > 0 ? o->inner_eps : 0.1;
    h->max_gap = o->inner_maxit > 0 ? o->inner_maxit : 500;
    h->ru_req = -1;
    h->ru_at = 0;
    h->restart = o->inner_restart ? 1 : 0;   /* 0: a refresh keeps the direction */
    if (ugk_h2d(w->st, h, sizeof(*h), stream)) return gpu_fail("state");
  }
  if (N && ufem_net_prepare(N, stream)) return AF_ERR;
  if (cc && ufem_coarse_build(w->cc, stream) != AF_OK) return AF_ERR;
  if (m == 1) {
    c.fused = N && !getenv("UFEM_NO_FUSED_C0T") && ufem_net_last_stage(N, &c.last);
    /* FP64 inference: asked for, or by default without the coarse term */
    c.f64 = c.fused && ufem_net_f64_ok(N) && (o->net_prec == 1 || (o->net_prec == 0 && !cc));
    if (c.f64 && ((rv = ufem_net_prepare64(N, stream)) != AF_OK || !ufem_net_last64(N, &c.last)))
      return rv != AF_OK ? rv : AF_ERR;
    if (ugk_f1_start(K, n, b, x, w->r, w->dinv, w->st, w->part, stream) ||
        (N && ugk_f1_input0(n, w->r, w->dinv, c.f64 ? NULL : ufem_net_input(N), c.f64 ? ufem_net_input64(N) : NULL,
                            ufem_net_mask(N), w->st, w->part, stream)))
      return gpu_fail("start");
    rv = run_chunks(w, stream, maxit, o->chunk > 0 ? o->chunk : 8, fcg1_step, &c, 1);
  } else {
    if (ugk_fcg_start(K, n, b, x, w->r, w->dinv, w->part, stream)) return gpu_fail("start");
    rv = run_chunks(w, stream, maxit, o->chunk > 0 ? o->chunk : 8, fcg_step, &c, 1);
  }
  if (rv != AF_OK) return rv;
  if (get_state(w, stream)) return AF_ERR;
  if (h->bb <= 0.0) {
    if (ugk_zero(x, sizeof(double) * (size_t)n, stream)) return gpu_fail("zero x");
    info->status = AF_OK;
    return AF_OK;
  }
  info->iters = h->iter;
  info->fallback = h->fb;
  info->outer = mixed ? h->n_ru : (h->fb ? h->fb_iter : 0);
  info->relres = sqrt(h->rr / h->bb);
  info->status = status_of(h->done);
  return info->status;
}

/* CG with the SSOR preconditioner (femEqsCGwSSOR, fem/fem_eqs.c:999-1188)
 * in colour order and in Eisenstat's form: b, x moved into colour order,
 * CG on the split-preconditioned system (one backward and one forward sweep
 * per iteration, no product with K), the original residual formed and
 * tested every SSOR_CHECK iterations, x recovered and moved back.  Uses the
 * workspace's BiCGStab vectors: s = b, t = x, rhat = x^, v = s, z = t-sweep,
 * shat = sqrt(diag), phat = the x-sweep. */
#define SSOR_CHECK 4
typedef struct {
  ufem_ssor *S;
  long n;
  int rule;
  ufem_work *w;
} ssor_ctx;

static int ssor_step(void *c, long it, void *stream)
{
  ssor_ctx *q = (ssor_ctx *)c;
  ufem_work *w = q->w;
  return ugk_eis_iter(&q->S->Kc, q->S->d_dpos, q->S->ncol, q->S->cb, q->n, w->rhat, w->r, w->p[0], w->q, w->z,
                      w->v, w->shat, w->t, w->phat, (it + 1) % SSOR_CHECK == 0, q->rule, w->st, w->part, stream);
}

static int solve_ssor(const ugk_mat *K, long n, const double *b, double *x, ufem_work *w, void *stream,
                      const ufem_solve_opts *o, long maxit, ufem_solve_info *info)
{
  ssor_ctx c;
  ufem_ssor *S;
  ugk_state *h = w->st_host;
  int rv, chunk = o->chunk > 0 ? o->chunk : 8;
  if (!w->v || !w->s || !w->t || !w->shat || !w->rhat || !w->phat) {
    ufem_set_error("SSOR: workspace without the BiCGStab vectors");
    return AF_ERR_VAL;
  }
  chunk = (chunk + SSOR_CHECK - 1) / SSOR_CHECK * SSOR_CHECK;   /* the check pattern repeats per chunk */
  rv = ufem_ssor_prepare(K, &w->ssor, stream);
  if (rv != AF_OK) return rv;
  S = w->ssor;
  rv = prepare(K, w, stream, o->eps, o->rule, maxit);
  if (rv != AF_OK) return rv;
  if (ugk_ssor_permute_vec(K->n_brows, K->bs, S->d_perm, b, w->s, 0, stream) ||
      ugk_ssor_permute_vec(K->n_brows, K->bs, S->d_perm, x, w->t, 0, stream))
    return gpu_fail("SSOR: colour order");
  if (ugk_fcg_start(&S->Kc, n, w->s, w->t, w->r, w->dinv, w->part, stream) ||
      ugk_dot(n, w->s, w->s, &w->st->bb, w->part, stream) ||
      ugk_eis_start(&S->Kc, S->d_dpos, S->ncol, S->cb, n, w->r, w->r, w->p[0], w->rhat, w->v, w->shat, w->st,
                    w->part, stream))
    return gpu_fail("SSOR: start");
  c.S = S;
  c.n = n;
  c.rule = o->rule;
  c.w = w;
  rv = run_chunks(w, stream, maxit, chunk, ssor_step, &c, 1);
  if (rv != AF_OK) return rv;
  if (ugk_eis_finish(&S->Kc, S->d_dpos, S->ncol, S->cb, w->rhat, w->t, w->phat, w->st, w->part, stream) ||
      ugk_ssor_permute_vec(K->n_brows, K->bs, S->d_perm, w->t, x, 1, stream))
    return gpu_fail("SSOR: order back");
  if (get_state(w, stream)) return AF_ERR;
  if (h->err == 4) { ufem_set_error("SSOR: a diagonal entry is not positive"); return AF_ERR_VAL; }
  if (h->bb <= 0.0) {
    if (ugk_zero(x, sizeof(double) * (size_t)n, stream)) return gpu_fail("zero x");
    info->status = AF_OK;
    return AF_OK;
  }
  info->iters = h->iter;
  info->outer = S->ncol;
  info->relres = sqrt(h->rr / h->bb);
  info->status = status_of(h->done);
  return info->status;
}

/* PCG with the two-level preconditioner z = dinv r + z_c (ugk_cc_iter): the
 * coarse matrix P^T K P is assembled and inverted first, on the device and
 * inside the timed solve, since it follows the current moduli */
typedef struct {
  const ugk_mat *K;
  long n;
  const double *b;
  double *x;
  ufem_work *w;
  int mixed;
} cc_ctx;

static int cc_step(void *c, long it, void *stream)
{
  cc_ctx *q = (cc_ctx *)c;
  ufem_work *w = q->w;
  (void)it;
  if (q->mixed)
    return ugk_cc_iter_mixed(q->K, w->cc, q->n, q->b, q->x, w->r, w->z, w->p[0], w->q, w->dinv, w->st, w->part,
                             stream);
  return ugk_cc_iter(q->K, w->cc, q->n, q->x, w->r, w->z, w->p[0], w->q, w->dinv, w->st, w->part, stream);
}

static int solve_twolevel(ugk_mat *Km, long n, const double *b, double *x, ufem_work *w, void *stream,
                          const ufem_solve_opts *o, long maxit, int mixed, ufem_solve_info *info)
{
  const ugk_mat *K = Km;
  cc_ctx c = {K, n, b, x, w, mixed};
  ugk_state *h = w->st_host;
  int rv;
  if (!w->cc) { ufem_set_error("two-level solver: no coarse grid (ufem_sys_set_coarse)"); return AF_ERR_VAL; }
  if (mixed) {
    if ((rv = ufem_mat_f32(Km, o->f32_rem, stream)) != AF_OK) return rv;
  }
  if ((rv = ufem_coarse_build(w->cc, stream)) != AF_OK) return rv;
  rv = prepare(K, w, stream, o->eps, o->rule, maxit);
  if (rv != AF_OK) return rv;
  if (mixed) {   /* the flexible-CG head's refresh rule */
    h->delta = o->inner_eps > 0 ? o->inner_eps : 0.1;
    h->max_gap = o->inner_maxit > 0 ? o->inner_maxit : 500;
    h->ru_req = -1;
    h->ru_at = 0;
    if (ugk_h2d(w->st, h, sizeof(*h), stream)) return gpu_fail("state");
  }
  if (ugk_fcg_start(K, n, b, x, w->r, w->dinv, w->part, stream)) return gpu_fail("start");
  rv = run_chunks(w, stream, maxit, o->chunk > 0 ? o->chunk : 8, cc_step, &c, 1);
  if (rv != AF_OK) return rv;
  if (get_state(w, stream)) return AF_ERR;
  if (h->bb <= 0.0) {
    if (ugk_zero(x, sizeof(double) * (size_t)n, stream)) return gpu_fail("zero x");
    info->status = AF_OK;
    return AF_OK;
  }
  info->iters = h->iter;
  info->outer = mixed ? h->n_ru : w->cc->nc;
  info->relres = sqrt(h->rr / h->bb);
  info->status = status_of(h->done);
  return info->status;
}

/* Mixed precision with FP64 vectors: Jacobi PCG whose only FP32 data is the
 * stored copy of K, widened to FP64 as it is loaded for q = K p; every
 * vector, dot product and recurrence stays FP64.  r drifts from b - K x as
 * the FP32 products accumulate, so r is recomputed in FP64 whenever the
 * recursive residual has fallen by inner_eps (default 0.1) since the last
 * recomputation, or after inner_maxit (default 500) iterations, and a
 * convergence is accepted on a recomputed residual only.  The refresh is
 * launched once per chunk of iterations (a request waits at most a chunk):
 * a launch that finds nothing to do still costs about 5 us. */
typedef struct {
  const ugk_mat *K;
  long n;
  const double *b;
  double *x;
  ufem_work *w;
  int chunk;    /* iterations per chunk; the refresh opens each */
} mx64_ctx;

static int mx64_step(void *c, long it, void *stream)
{
  mx64_ctx *m = (mx64_ctx *)c;
  ufem_work *w = m->w;
  return ugk_mx64_iter(m->K, m->n, m->b, m->x, w->r, w->p[0], w->q, w->dinv, w->st, w->part, it % m->chunk == 0,
                       stream);
}

static int solve_mixed64(ugk_mat *Km, long n, const double *b, double *x, ufem_work *w, void *stream,
                         const ufem_solve_opts *o, long maxit, ufem_solve_info *info)
{
  const ugk_mat *K = Km;
  const int chunk = o->chunk > 0 ? o->chunk : 16;
  mx64_ctx c = {K, n, b, x, w, chunk < 2 ? 2 : chunk + (chunk & 1)};   /* as run_chunks rounds it */
  ugk_state *h = w->st_host;
  int rv;
  if ((rv = ufem_mat_f32(Km, o->f32_rem, stream)) != AF_OK) return rv;
  rv = prepare(K, w, stream, o->eps, o->rule, maxit);
  if (rv != AF_OK) return rv;
  h->delta = o->inner_eps > 0 ? o->inner_eps : 0.1;
  h->max_gap = o->inner_maxit > 0 ? o->inner_maxit : 500;
  h->ru_req = -1;
  h->ru_at = 0;
  if (ugk_h2d(w->st, h, sizeof(*h), stream)) return gpu_fail("state");
  if (ugk_fcg_start(K, n, b, x, w->r, w->dinv, w->part, stream)) return gpu_fail("start");
  rv = run_chunks(w, stream, maxit, chunk, mx64_step, &c, 1);
  if (rv != AF_OK) return rv;
  if (get_state(w, stream)) return AF_ERR;
  if (h->bb <= 0.0) {
    if (ugk_zero(x, sizeof(double) * (size_t)n, stream)) return gpu_fail("zero x");
    info->status = AF_OK;
    return AF_OK;
  }
  info->iters = h->iter;
  info->outer = h->n_ru;
  info->relres = sqrt(h->rr / h->bb);
  info->status = status_of(h->done);
  return info->status;
}

/* Jacobi PCG (the streamed single-reduction form, kernels launched one by
 * one, no graph) with every vector, partial sum and scalar it needs beyond
 * b and x in the fixed device workspace of ugk_noalloc_workspace: the call
 * allocates nothing -- no device or pinned memory, no events, no graph.
 * The state is read back with a plain copy after every chunk of iterations
 * and the time is taken on the host clock.  Calls on one GPU share its
 * workspace: they must not overlap. */
int ufem_solve_device_noalloc(const ugk_mat *K, long n, const double *b, double *x, void *stream,
                              const ufem_solve_opts *o, ufem_solve_info *info)
{
  static ugk_state hbuf;   /* host copy of the workspace's state (same race rule) */
  const long N = UGK_NOALLOC_MAX_N;
  const int chunk = o->chunk > 0 ? o->chunk : 32;
  long maxit = o->maxit > 0 ? o->maxit : (o->maxit == 0 ? n : 0), it = 0, k;
  void *base = NULL;
  size_t bytes = 0;
  double *d, t0, bb;
  struct timespec ts;
  ufem_work w;
  int rv;
  if (o->solver != UFEM_SOLVER_PCG || o->exec != UFEM_EXEC_STREAMED) {
    ufem_set_error("no-alloc solve: Jacobi PCG in the streamed form only");
    return AF_ERR_TYP;
  }
  if (n > N) {
    ufem_set_error("no-alloc solve: %ld unknowns, the fixed workspace holds at most %ld", n, N);
    return AF_ERR_BIG;
  }
  if (sizeof(ugk_state) > sizeof(double) * UGK_NOALLOC_STATE_DOUBLES) {
    ufem_set_error("no-alloc solve: solver state larger than its slot");
    return AF_ERR;
  }
  if (ugk_noalloc_workspace(&base, &bytes)) return gpu_fail("no-alloc workspace");
  d = (double *)base;
  memset(&w, 0, sizeof(w));
  w.n = n;
  w.r = d;
  w.z = d + N;
  w.p[0] = d + 2 * N;
  w.p[1] = d + 3 * N;
  w.q = d + 4 * N;
  w.dinv = d + 5 * N;
  w.part = d + 6 * N;
  w.st = (ugk_state *)(d + 6 * N + UGK_PART_DOUBLES);
  w.st_host = &hbuf;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  t0 = 1e3 * (double)ts.tv_sec + 1e-6 * (double)ts.tv_nsec;
  memset(info, 0, sizeof(*info));
  rv = prepare(K, &w, stream, o->eps, o->rule, maxit);
  if (rv != AF_OK) return rv;
  if (ugk_cg1_start(K, n, b, x, w.r, w.z, w.p[0], w.p[1], w.dinv, w.part, stream)) return gpu_fail("start");
  for (;;) {
    for (k = 0; k < chunk; k++)
      if (ugk_cg1_iter(K, n, x, w.r, w.z, w.p[0], w.p[1], w.q, w.dinv, w.st, w.part, stream))
        return gpu_fail("iteration");
    it += chunk;
    if (get_state(&w, stream)) return AF_ERR;
    if (hbuf.done || it > maxit + 2L * chunk) break;
  }
  if (hbuf.bb <= 0.0) {   /* zero load: x = 0 */
    if (ugk_zero(x, sizeof(double) * (size_t)n, stream)) return gpu_fail("zero x");
    info->status = AF_OK;
  } else {
    info->iters = hbuf.iter;
    info->relres = sqrt(hbuf.rr / hbuf.bb);
    info->status = status_of(hbuf.done);
  }
  if (ugk_dot(n, b, b, &w.st->bb, w.part, stream) || get_state(&w, stream)) return gpu_fail("norm b");
  bb = hbuf.bb;
  if (true_residual(K, n, b, x, &w, stream, bb, &info->true_relres)) return AF_ERR;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  info->solve_ms = 1e3 * (double)ts.tv_sec + 1e-6 * (double)ts.tv_nsec - t0;
  return info->status;
}

int ufem_solve_device(const ugk_mat *K, long n, const double *b, double *x, ufem_work *w, void *stream,
                      void *ev0, void *ev1, const ufem_solve_opts *o, ufem_solve_info *info)
{
  long maxit = o->maxit > 0 ? o->maxit : (o->maxit == 0 ? n : 0);
  int rv;
  double bb;
  if (ugk_event_record(ev0, stream)) return gpu_fail("event");
  switch (o->solver) {
    case UFEM_SOLVER_PCG: rv = solve_pcg(K, n, b, x, w, stream, o, maxit, info); break;
    case UFEM_SOLVER_BICGSTAB: rv = solve_bicgs(K, n, b, x, w, stream, o, maxit, info); break;
    case UFEM_SOLVER_MIXED: rv = solve_mixed((ugk_mat *)K, n, b, x, w, stream, o, maxit, info); break;
    case UFEM_SOLVER_LEARNED: rv = solve_learned((ugk_mat *)K, n, b, x, w, stream, o, maxit, 0, info); break;
    case UFEM_SOLVER_LEARNED_MIXED:
      rv = solve_learned((ugk_mat *)K, n, b, x, w, stream, o, maxit, 1, info);
      break;
    case UFEM_SOLVER_SSOR: rv = solve_ssor(K, n, b, x, w, stream, o, maxit, info); break;
    case UFEM_SOLVER_TWOLEVEL: rv = solve_twolevel((ugk_mat *)K, n, b, x, w, stream, o, maxit, 0, info); break;
    case UFEM_SOLVER_TWOLEVEL_MIXED:
      rv = solve_twolevel((ugk_mat *)K, n, b, x, w, stream, o, maxit, 1, info);
      break;
    case UFEM_SOLVER_MIXED_V64: rv = solve_mixed64((ugk_mat *)K, n, b, x, w, stream, o, maxit, info); break;
    default: ufem_set_error("unknown solver %d", o->solver); return AF_ERR_VAL;
  }
  if (ugk_event_record(ev1, stream)) return gpu_fail("event");
  info->solve_ms = ugk_event_ms(ev0, ev1);
  if (rv == AF_ERR_ZER || rv == AF_ERR_MEM) return rv;
  if (ugk_dot(n, b, b, &w->st->bb, w->part, stream) || get_state(w, stream)) return gpu_fail("norm b");
  bb = w->st_host->bb;
  if (true_residual(K, n, b, x, w, stream, bb, &info->true_relres)) return AF_ERR;
  return rv;
}
