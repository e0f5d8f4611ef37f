/*
 * ufem_net.c -- the learned preconditioner: a symmetric L-level 3-D
 * convolution U-Net (L = 2..8, 3 by default) on the structured node grid of
 * a box mesh, trained and applied on the GPU by the kernels of ugk_unet.hip.
 *
 * The network has no counterpart in the reference; it accelerates the
 * Jacobi-PCG solve (femEqsCGwJ, fem/fem_eqs.c:88-237) as the preconditioner
 *
 *   z = dinv r + mask dsq N(mask dsq r),   dsq = sqrt(dinv),
 *   N(v) = C0^T (g0 h0 + C1^T (g1 h1 + ... + C_{L-1}^T (g_{L-1} h_{L-1}))),
 *   h0 = C0 v, h_l = C_l h_{l-1},
 *   g_l = exp(b_l - (1 + a_l) ebar_l)   (per node and channel),
 *
 * with C0 a 3x3x3 convolution 3 -> 16 channels, C1 16 -> 32, C2 32 -> 64
 * and C3.. 64 -> 64 stride-2 convolutions (levels of (n + 1) / 2 nodes per
 * axis), C^T their adjoints (the same weights, transposed: the up path is
 * tied to the down path), and ebar_l the log-modulus log(E / 30 GPa) at the
 * nodes, averaged over 3x3x3 windows per coarser level.  More levels widen
 * the receptive field towards the beam's bending modes.  N = sum_l H_l^T
 * diag(g_l) H_l is
 * linear in r, symmetric and positive semi-definite, so M is SPD and the
 * short recurrences of CG stay valid.  The input is scaled by 1 / rms(dsq r)
 * and the output by rms(dsq r) only to keep the numbers near 1.
 *
 * Weights are bf16 in the matrix cores (f32 masters for Adam); activations
 * are f32 and enter the bf16 matrix cores as exact three-way splits, so the
 * preconditioner is linear to f32 rounding (bf16 activations would add
 * noise that costs CG many iterations).
 *
 * Training pairs come from plain Jacobi-PCG solves on random material
 * fields: snapshots (x_k, r_k) every few iterations, and once the solve has
 * converged to u the error e_k = u - x_k, so that K e_k = r_k.  The loss of
 * a pair is the energy-norm error of z relative to that of e,
 * L = (e - z)^T (r - K z) / (e^T r), whose gradient in z is -2 (r - K z) /
 * (e^T r): one FP32 SpMV with that field's matrix per pair and step.
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ufem_fem_math.h"
#include "ufem_internal.h"

#define E_REF 30e9

#define NL_MAX 8
/* channels of the node-grid form: output and input of C0, C1, ..; the
 * coarse-first form drops C0 and starts with a stride-2 3 -> 32 layer */
static const int CH_F[NL_MAX + 1] = {16, 32, 64, 64, 64, 64, 64, 64, 64};
static const int CI_F[NL_MAX + 1] = {3, 16, 32, 64, 64, 64, 64, 64, 64};

struct ufem_net {
  int L;                          /* levels */
  int fine;                       /* 1: C0 is stride 1 on the node grid; 0: coarse first */
  int ch[NL_MAX], ci[NL_MAX];     /* output and input channels of C_l */
  ugk_grid3 gf;                   /* the node grid (network input and output) */
  long Vf;
  ugk_grid3 g[NL_MAX];
  long V[NL_MAX], Vt;             /* nodes per level, and their sum */
  int nbmax;
  void *stream;
  void *ev[2];
  long P, offW[NL_MAX], offA[NL_MAX], offB[NL_MAX];
  float *w, *gr, *m, *v;          /* packed f32 parameters, gradient, Adam moments */
  unsigned short *wf[NL_MAX], *wt[NL_MAX];  /* bf16 [27][cop][cip] and [27][cip][cop] */
  float *dwpart;
  double *gpart;
  /* activations (f32, nbmax samples) */
  float *x0, *h[NL_MAX], *t[NL_MAX], *gt[NL_MAX], *eb[NL_MAX], *y;
  float *dy, *dt[NL_MAX], *dh[NL_MAX];
  float *zz, *kz;
  double *dpart, *hpart;
  long step;
  /* training pairs */
  long cap, count;
  float *pv, *prr, *pee;
  double *per;
  int *pfield;
  /* material fields of the pairs: FP32 K values, dinv, ebar levels */
  int nfield, capf;
  float **fvals, **fdinv, **feb;
  /* binding to a system */
  ufem_sys *sys;
  float *mask;
  float *yi, *gi[NL_MAX], *ebi;
  unsigned short *x0s;            /* inference input, split format [V][3][4] */
  float *ebh;                     /* host ebar levels of the bound field */
  double *snap;
  long snap_cap, snap_n;
  int snap_every;
  /* FP64 inference (the solver's path): activations [V][C], weight copies */
  double *h64[NL_MAX], *t64[NL_MAX], *fd64[NL_MAX], *fu64[NL_MAX], *wu64;
  double *v64, *y64;
};

static int r16(int c) { return (c + 15) / 16 * 16; }

static int net_fail(const char *what)
{
  ufem_set_error("%s: %s", what, ugk_error());
  return AF_ERR;
}

static ugk_grid3 coarse(ugk_grid3 g)
{
  ugk_grid3 c = {(g.D + 1) / 2, (g.H + 1) / 2, (g.W + 1) / 2};
  return c;
}

/* channels of C_l, capped at `width` (a multiple of 4, 4 .. 64) */
static void channels(int levels, int fine, int width, int *ch, int *ci)
{
  int l;
  for (l = 0; l < levels; l++) {
    ch[l] = fine ? CH_F[l] : CH_F[l + 1];
    ci[l] = fine ? CI_F[l] : (l ? CH_F[l] : 3);
    if (ch[l] > width) ch[l] = width;
    if (ci[l] > width) ci[l] = width;
  }
}

static int width_ok(int width) { return width >= 4 && width <= 64 && width % 4 == 0; }

long ufem_net_nparams_width(int levels, int fine, int width)
{
  int ch[NL_MAX], ci[NL_MAX], l;
  long P = 0;
  if (levels < 2 || levels > NL_MAX || !width_ok(width)) return -1;
  channels(levels, fine, width, ch, ci);
  for (l = 0; l < levels; l++) P += 27L * ci[l] * ch[l] + 2L * ch[l];
  return P;
}

long ufem_net_nparams_form(int levels, int fine) { return ufem_net_nparams_width(levels, fine, 64); }

long ufem_net_nparams_levels(int levels) { return ufem_net_nparams_form(levels, 1); }
long ufem_net_nparams(void) { return ufem_net_nparams_levels(3); }
int ufem_net_fine(const ufem_net *N) { return N->fine; }

/* grid of level lev (-1: the node grid); stride of C_l and the level it reads */
static ugk_grid3 grid_of(const ufem_net *N, int lev) { return lev < 0 ? N->gf : N->g[lev]; }
static int stride_of(const ufem_net *N, int l) { return (l == 0 && N->fine) ? 1 : 2; }
static int lev_in_of(const ufem_net *N, int l) { return (l == 0 && N->fine) ? 0 : l - 1; }
/* the level of the network's input and output: the node grid */
static int io_lev(const ufem_net *N) { return N->fine ? 0 : -1; }
int ufem_net_levels(const ufem_net *N) { return N->L; }

static void *dalloc(size_t bytes, int *bad)
{
  void *p = ugk_malloc(bytes);
  if (!p) *bad = 1;
  return p;
}

void ufem_net_destroy(ufem_net *N)
{
  int l, f;
  if (!N) return;
  if (N->stream) (void)ugk_sync(N->stream);
  ugk_free(N->w); ugk_free(N->gr); ugk_free(N->m); ugk_free(N->v);
  for (l = 0; l < NL_MAX; l++) {
    ugk_free(N->wf[l]); ugk_free(N->wt[l]);
    ugk_free(N->h[l]); ugk_free(N->t[l]); ugk_free(N->gt[l]); ugk_free(N->eb[l]);
    ugk_free(N->dt[l]); ugk_free(N->dh[l]); ugk_free(N->gi[l]);
  }
  ugk_free(N->dwpart); ugk_free(N->gpart);
  ugk_free(N->x0); ugk_free(N->y); ugk_free(N->dy); ugk_free(N->zz); ugk_free(N->kz); ugk_free(N->dpart);
  ugk_free(N->pv); ugk_free(N->prr); ugk_free(N->pee);
  for (f = 0; f < N->nfield; f++) { ugk_free(N->fvals[f]); ugk_free(N->fdinv[f]); ugk_free(N->feb[f]); }
  free(N->fvals); free(N->fdinv); free(N->feb);
  ugk_free(N->mask); ugk_free(N->yi); ugk_free(N->ebi); ugk_free(N->snap);
  ugk_free(N->x0s);
  for (l = 0; l < NL_MAX; l++) { ugk_free(N->h64[l]); ugk_free(N->t64[l]); ugk_free(N->fd64[l]); ugk_free(N->fu64[l]); }
  ugk_free(N->wu64); ugk_free(N->v64); ugk_free(N->y64);
  free(N->hpart); free(N->per); free(N->pfield); free(N->ebh);
  if (N->ev[0]) ugk_event_destroy(N->ev[0]);
  if (N->ev[1]) ugk_event_destroy(N->ev[1]);
  if (N->stream) ugk_stream_destroy(N->stream);
  free(N);
}

static ugk_dw_desc dw_desc(const ufem_net *N, int l, const float *x, int ldx, int cin, const float *dy, int ldd,
                           int cout, int stride, int lev_in, int nb)
{
  ugk_dw_desc d;
  memset(&d, 0, sizeof(d));
  d.x = x; d.ldx = ldx; d.cin = cin; d.x_f32 = 1;
  d.dy = dy; d.ldd = ldd; d.cout = cout; d.dy_f32 = 1;
  d.cip = r16(cin); d.cop = r16(cout);
  d.mode = stride == 2 ? UGK_CONV_S2 : UGK_CONV_S1;
  d.nb = nb;
  d.gi = grid_of(N, lev_in);
  d.go = grid_of(N, stride == 2 ? lev_in + 1 : lev_in);
  (void)l;
  return d;
}

ufem_net *ufem_net_create(int W, int H, int D, int batch)
{
  return ufem_net_create_levels(W, H, D, batch, 3);
}

ufem_net *ufem_net_create_levels(int W, int H, int D, int batch, int levels)
{
  return ufem_net_create_form(W, H, D, batch, levels, 1);
}

ufem_net *ufem_net_create_form(int W, int H, int D, int batch, int levels, int fine)
{
  return ufem_net_create_width(W, H, D, batch, levels, fine, 64);
}

ufem_net *ufem_net_create_width(int W, int H, int D, int batch, int levels, int fine, int width)
{
  ufem_net *N;
  int l, bad = 0;
  long nb, maxpart = 0, n;
  if (W < 2 || H < 2 || D < 2 || batch < 1) { ufem_set_error("net grid or batch out of range"); return NULL; }
  if (levels < 2 || levels > NL_MAX) { ufem_set_error("levels %d out of range (2..%d)", levels, NL_MAX); return NULL; }
  if (!width_ok(width)) { ufem_set_error("width %d out of range (a multiple of 4, 4..64)", width); return NULL; }
  if (ugk_init()) { net_fail("device"); return NULL; }
  N = (ufem_net *)calloc(1, sizeof(*N));
  if (!N) return NULL;
  N->L = levels;
  N->fine = fine ? 1 : 0;
  channels(levels, N->fine, width, N->ch, N->ci);
  N->gf.D = D; N->gf.H = H; N->gf.W = W;
  N->Vf = (long)D * H * W;
  N->g[0] = N->fine ? N->gf : coarse(N->gf);
  for (l = 1; l < levels; l++) N->g[l] = coarse(N->g[l - 1]);
  N->Vt = 0;
  for (l = 0; l < levels; l++) {
    N->V[l] = (long)N->g[l].D * N->g[l].H * N->g[l].W;
    N->Vt += N->V[l];
  }
  N->nbmax = batch;
  nb = batch;
  n = 3 * N->Vf;
  N->stream = ugk_stream_create();
  N->ev[0] = ugk_event_create();
  N->ev[1] = ugk_event_create();
  if (!N->stream || !N->ev[0] || !N->ev[1]) { net_fail("stream"); ufem_net_destroy(N); return NULL; }
  N->P = 0;
  for (l = 0; l < levels; l++) { N->offW[l] = N->P; N->P += 27L * N->ci[l] * N->ch[l]; }
  for (l = 0; l < levels; l++) {
    N->offA[l] = N->P; N->P += N->ch[l];
    N->offB[l] = N->P; N->P += N->ch[l];
  }
  N->w = (float *)dalloc(sizeof(float) * N->P, &bad);
  N->gr = (float *)dalloc(sizeof(float) * N->P, &bad);
  N->m = (float *)dalloc(sizeof(float) * N->P, &bad);
  N->v = (float *)dalloc(sizeof(float) * N->P, &bad);
  for (l = 0; l < levels; l++) {
    const size_t b = sizeof(unsigned short) * 27 * r16(N->ci[l]) * r16(N->ch[l]);
    const long Vl = N->V[l];
    N->wf[l] = (unsigned short *)dalloc(b, &bad);
    N->wt[l] = (unsigned short *)dalloc(b, &bad);
    N->h[l] = (float *)dalloc(sizeof(float) * (size_t)(nb * Vl * N->ch[l]), &bad);
    N->t[l] = (float *)dalloc(sizeof(float) * (size_t)(nb * Vl * N->ch[l]), &bad);
    N->gt[l] = (float *)dalloc(sizeof(float) * (size_t)(nb * Vl * N->ch[l]), &bad);
    N->dt[l] = (float *)dalloc(sizeof(float) * (size_t)(nb * Vl * N->ch[l]), &bad);
    N->dh[l] = (float *)dalloc(sizeof(float) * (size_t)(nb * Vl * N->ch[l]), &bad);
    N->eb[l] = (float *)dalloc(sizeof(float) * (size_t)(nb * Vl), &bad);
    N->gi[l] = (float *)dalloc(sizeof(float) * (size_t)(Vl * N->ch[l]), &bad);
    {
      /* weight-gradient partials: the largest of the calls the backward makes */
      ugk_dw_desc d = dw_desc(N, l, NULL, 4, N->ci[l], NULL, 4, N->ch[l], stride_of(N, l), lev_in_of(N, l), batch);
      long need = (long)ugk_conv_dw_slices(&d) * 27L * d.cop * d.cip;
      if (need > maxpart) maxpart = need;
    }
  }
  N->dwpart = (float *)dalloc(sizeof(float) * maxpart, &bad);
  N->gpart = (double *)dalloc(sizeof(double) * 512 * 2 * 64, &bad);
  N->x0 = (float *)dalloc(sizeof(float) * (size_t)(nb * N->Vf * 4), &bad);
  N->y = (float *)dalloc(sizeof(float) * (size_t)(nb * N->Vf * 4), &bad);
  N->dy = (float *)dalloc(sizeof(float) * (size_t)(nb * N->Vf * 4), &bad);
  N->zz = (float *)dalloc(sizeof(float) * (size_t)n, &bad);
  N->kz = (float *)dalloc(sizeof(float) * (size_t)n, &bad);
  N->dpart = (double *)dalloc(sizeof(double) * 1024 * (size_t)(nb > 4 ? nb : 4), &bad);
  N->hpart = (double *)malloc(sizeof(double) * 1024 * (size_t)(nb > 4 ? nb : 4));
  N->yi = (float *)dalloc(sizeof(float) * (size_t)(N->Vf * 4), &bad);
  N->x0s = (unsigned short *)dalloc(3 * sizeof(unsigned short) * (size_t)(N->Vf * 4), &bad);
  N->ebi = (float *)dalloc(sizeof(float) * (size_t)N->Vt, &bad);
  if (bad || !N->hpart) { net_fail("network buffers"); ufem_net_destroy(N); return NULL; }
  return N;
}

/* ---------------- weights ---------------------------------------------- */
static int adam_all(ufem_net *N, float lr)
{
  int l;
  if (lr > 0) N->step++;
  for (l = 0; l < N->L; l++) {
    if (ugk_adam(27L * N->ci[l] * N->ch[l], N->ch[l], N->ci[l], r16(N->ch[l]), r16(N->ci[l]), N->w + N->offW[l], N->gr + N->offW[l],
                 N->m + N->offW[l], N->v + N->offW[l], N->wf[l], N->wt[l], lr, 0.9f, 0.999f, 1e-8f, N->step,
                 N->stream))
      return net_fail("weights");
    if (lr > 0 &&
        (ugk_adam(N->ch[l], 1, N->ch[l], 0, 0, N->w + N->offA[l], N->gr + N->offA[l], N->m + N->offA[l],
                  N->v + N->offA[l], NULL, NULL, lr, 0.9f, 0.999f, 1e-8f, N->step, N->stream) ||
         ugk_adam(N->ch[l], 1, N->ch[l], 0, 0, N->w + N->offB[l], N->gr + N->offB[l], N->m + N->offB[l],
                  N->v + N->offB[l], NULL, NULL, lr, 0.9f, 0.999f, 1e-8f, N->step, N->stream)))
      return net_fail("gate parameters");
  }
  return AF_OK;
}

int ufem_net_set_weights(ufem_net *N, const float *w)
{
  if (ugk_h2d(N->w, w, sizeof(float) * N->P, N->stream)) return net_fail("weights upload");
  if (ugk_zero(N->m, sizeof(float) * N->P, N->stream) || ugk_zero(N->v, sizeof(float) * N->P, N->stream))
    return net_fail("moments");
  N->step = 0;
  if (adam_all(N, 0.f)) return AF_ERR;
  return ugk_sync(N->stream) ? net_fail("sync") : AF_OK;
}

int ufem_net_get_weights(ufem_net *N, float *w)
{
  return ugk_d2h(w, N->w, sizeof(float) * N->P, N->stream) ? net_fail("weights download") : AF_OK;
}

static unsigned long long mix64(unsigned long long z)
{
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

/* weights N(0, 1 / (27 cin)) from a counter-based generator; gates a = 0,
 * b = -3, so that the untrained correction is small (g ~ 0.05 / E-ratio) and
 * training starts close to Jacobi PCG */
int ufem_net_init(ufem_net *N, unsigned long seed)
{
  float *w = (float *)calloc((size_t)N->P, sizeof(float));
  long i = 0;
  int l, rv;
  if (!w) return AF_ERR_MEM;
  for (l = 0; l < N->L; l++) {
    long k;
    for (k = 0; k < N->ch[l]; k++) w[N->offB[l] + k] = -3.f;
  }
  for (l = 0; l < N->L; l++) {
    const long Pl = 27L * N->ci[l] * N->ch[l];
    const double sd = 1.0 / sqrt(27.0 * N->ci[l]);
    long k;
    for (k = 0; k < Pl; k++, i++) {
      unsigned long long a = mix64((unsigned long long)seed * 0x632BE59BD9B4E019ULL + 2ULL * (unsigned long long)i);
      unsigned long long b = mix64(a ^ 0xD1B54A32D192ED03ULL);
      double u1 = ((double)(a >> 11) + 0.5) / 9007199254740992.0;
      double u2 = (double)(b >> 11) / 9007199254740992.0;
      w[N->offW[l] + k] = (float)(sd * sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
    }
  }
  rv = ufem_net_set_weights(N, w);
  free(w);
  return rv;
}

/* ---------------- forward / backward ------------------------------------- */
/* layer l (or its adjoint) from level lev_in to lev_out, input format xf
 * (UGK_ACT_*), f32 output; epilogue (* emul) (+ ga gb) */
static int conv_f(ufem_net *N, const void *x, int xf, int ldx, int cin, int l, int adjoint, float *y, int ldy,
                  int cout, int lev_in, int lev_out, const float *emul, const float *ga, const float *gb, int ldg,
                  int nb, const int *skip, void *stream)
{
  ugk_conv_desc d;
  memset(&d, 0, sizeof(d));
  d.x = x; d.ldx = ldx; d.cin = cin; d.x_f32 = xf;
  d.w = adjoint ? N->wt[l] : N->wf[l];
  d.cip = adjoint ? r16(N->ch[l]) : r16(N->ci[l]);
  d.cop = adjoint ? r16(N->ci[l]) : r16(N->ch[l]);
  d.y = y; d.ldy = ldy; d.cout = cout; d.y_f32 = 1;
  d.emul = emul; d.ldm = ldg;
  d.ga = ga; d.ldga = ldg; d.gb = gb; d.ldgb = ldg;
  d.mode = stride_of(N, l) == 1 ? UGK_CONV_S1 : (adjoint ? UGK_CONV_T2 : UGK_CONV_S2);
  d.flip = (stride_of(N, l) == 1 && adjoint) ? 1 : 0;
  d.nb = nb;
  d.gi = grid_of(N, lev_in);
  d.go = grid_of(N, lev_out);
  d.skip = skip;
  return ugk_conv(&d, stream);
}

/* the same on f32 activations (training) */
static int conv(ufem_net *N, const float *x, int ldx, int cin, int l, int adjoint, float *y, int ldy, int cout,
                int lev_in, int lev_out, const float *ga, const float *gb, int ldg, int nb, const int *skip,
                void *stream)
{
  return conv_f(N, x, UGK_ACT_F32, ldx, cin, l, adjoint, y, ldy, cout, lev_in, lev_out, NULL, ga, gb, ldg, nb, skip,
                stream);
}

/* stage k of y = N(v) for nb samples with gates g[l] (nb x V_l x C_l), L
 * levels, 2 L + 1 stages: k = 0: h0 = C0 v; k = 1 .. L-1: h_k = C_k h_{k-1};
 * k = L: t_{L-1} = g_{L-1} h_{L-1}; k = L+1 .. 2L-1: t_l = g_l h_l +
 * C_{l+1}^T t_{l+1} with l = 2L-1-k; k = 2L: y = C0^T t0 */
static int net_stages(const ufem_net *N) { return 2 * N->L + 1; }
static int net_stage(ufem_net *N, int k, const float *v, float *const *g, float *y, int nb, const int *skip,
                     void *st)
{
  const int L = N->L;
  if (k == 0) return conv(N, v, 4, 4, 0, 0, N->h[0], N->ch[0], N->ch[0], io_lev(N), 0, NULL, NULL, 0, nb, skip, st);
  if (k < L)
    return conv(N, N->h[k - 1], N->ch[k - 1], N->ch[k - 1], k, 0, N->h[k], N->ch[k], N->ch[k], k - 1, k, NULL, NULL, 0, nb,
                skip, st);
  if (k == L) return ugk_mul((long)nb * N->V[L - 1] * N->ch[L - 1], g[L - 1], N->h[L - 1], N->t[L - 1], skip, st);
  if (k < 2 * L) {
    const int l = 2 * L - 1 - k;
    return conv(N, N->t[l + 1], N->ch[l + 1], N->ch[l + 1], l + 1, 1, N->t[l], N->ch[l], N->ch[l], l + 1, l, g[l], N->h[l],
                N->ch[l], nb, skip, st);
  }
  return conv(N, N->t[0], N->ch[0], N->ch[0], 0, 1, y, 4, 3, 0, io_lev(N), NULL, NULL, 0, nb, skip, st);
}

static int net_forward(ufem_net *N, const float *v, float *const *g, float *y, int nb, const int *skip,
                       void *st)
{
  int k;
  for (k = 0; k < net_stages(N); k++)
    if (net_stage(N, k, v, g, y, nb, skip, st)) return net_fail("network forward");
  return AF_OK;
}

/* inference, one sample, input x0s in the split format (made once by the
 * producer, so the first layer does no per-load split work), f32
 * activations in the first sample of the training buffers, 2 L stages:
 * k = 0: h0 = C0 v; k = 1 .. L-2: h_k = C_k h_{k-1}; k = L-1: t_{L-1} =
 * g_{L-1} C_{L-1} h_{L-2} (the gate product in the convolution's epilogue);
 * k = L .. 2L-2: t_l = g_l h_l + C_{l+1}^T t_{l+1} with l = 2L-2-k;
 * k = 2L-1: y = C0^T t0 (f32 [V][4]); the same products as net_forward,
 * bit for bit */
static int infer_stages(const ufem_net *N) { return 2 * N->L; }
static int infer_stage(ufem_net *N, int k, float *const *g, float *y, const int *skip, void *st)
{
  const int F = UGK_ACT_F32, L = N->L;
  if (k == 0)
    return conv_f(N, N->x0s, UGK_ACT_SPLIT, 4, 4, 0, 0, N->h[0], N->ch[0], N->ch[0], io_lev(N), 0, NULL, NULL, NULL, 0, 1,
                  skip, st);
  if (k < L - 1)
    return conv_f(N, N->h[k - 1], F, N->ch[k - 1], N->ch[k - 1], k, 0, N->h[k], N->ch[k], N->ch[k], k - 1, k, NULL, NULL, NULL,
                  0, 1, skip, st);
  if (k == L - 1)
    return conv_f(N, N->h[k - 1], F, N->ch[k - 1], N->ch[k - 1], k, 0, N->t[k], N->ch[k], N->ch[k], k - 1, k, g[k], NULL, NULL,
                  N->ch[k], 1, skip, st);
  if (k < 2 * L - 1) {
    const int l = 2 * L - 2 - k;
    return conv_f(N, N->t[l + 1], F, N->ch[l + 1], N->ch[l + 1], l + 1, 1, N->t[l], N->ch[l], N->ch[l], l + 1, l, NULL, g[l],
                  N->h[l], N->ch[l], 1, skip, st);
  }
  return conv_f(N, N->t[0], F, N->ch[0], N->ch[0], 0, 1, y, 4, 3, 0, io_lev(N), NULL, NULL, NULL, 0, 1, skip, st);
}

static int infer(ufem_net *N, float *const *g, float *y, const int *skip, void *st)
{
  int k;
  for (k = 0; k < infer_stages(N); k++)
    if (infer_stage(N, k, g, y, skip, st)) return net_fail("network inference");
  return AF_OK;
}

static int wgrad(ufem_net *N, int l, const float *x, int ldx, int cin, const float *dy, int ldd, int lev_in,
                 int nb, int acc)
{
  ugk_dw_desc d = dw_desc(N, l, x, ldx, cin, dy, ldd, N->ch[l], stride_of(N, l), lev_in, nb);
  return ugk_conv_dw(&d, N->dwpart, ugk_conv_dw_slices(&d), N->gr + N->offW[l], acc, N->stream);
}

/* parameter gradient of <dy, N(v)> into N->gr (forward fields in place):
 * dt0 = C0 dy, dt_l = C_l dt_{l-1} (the adjoints of the up path), with the
 * weight gradients of the up path's C_l^T; dh_{L-1} = g_{L-1} dt_{L-1},
 * dh_{l-1} = C_l^T dh_l + g_{l-1} dt_{l-1} (the down path), with the weight
 * gradients of its C_l; the gate gradients from h_l and dt_l */
static int net_backward(ufem_net *N, const float *v, const float *dy, int nb)
{
  float *const *g = N->gt;
  const int L = N->L;
  int l;
  if (conv(N, dy, 4, 4, 0, 0, N->dt[0], N->ch[0], N->ch[0], io_lev(N), 0, NULL, NULL, 0, nb, NULL, N->stream) ||
      wgrad(N, 0, dy, 4, 3, N->t[0], N->ch[0], io_lev(N), nb, 0))
    return net_fail("network backward");
  for (l = 1; l < L; l++)
    if (conv(N, N->dt[l - 1], N->ch[l - 1], N->ch[l - 1], l, 0, N->dt[l], N->ch[l], N->ch[l], l - 1, l, NULL, NULL, 0, nb, NULL,
             N->stream) ||
        wgrad(N, l, N->dt[l - 1], N->ch[l - 1], N->ch[l - 1], N->t[l], N->ch[l], l - 1, nb, 0))
      return net_fail("network backward");
  if (ugk_mul((long)nb * N->V[L - 1] * N->ch[L - 1], g[L - 1], N->dt[L - 1], N->dh[L - 1], NULL, N->stream))
    return net_fail("network backward");
  for (l = L - 1; l >= 1; l--)
    if (conv(N, N->dh[l], N->ch[l], N->ch[l], l, 1, N->dh[l - 1], N->ch[l - 1], N->ch[l - 1], l, l - 1, g[l - 1], N->dt[l - 1],
             N->ch[l - 1], nb, NULL, N->stream) ||
        wgrad(N, l, N->h[l - 1], N->ch[l - 1], N->ch[l - 1], N->dh[l], N->ch[l], l - 1, nb, 1))
      return net_fail("network backward");
  if (wgrad(N, 0, v, 4, 3, N->dh[0], N->ch[0], io_lev(N), nb, 1)) return net_fail("network backward");
  for (l = 0; l < L; l++)
    if (ugk_gate_grad((long)nb * N->V[l], N->ch[l], g[l], N->h[l], N->dt[l], N->eb[l], N->gpart,
                      N->gr + N->offA[l], N->gr + N->offB[l], N->stream))
      return net_fail("gate gradient");
  return AF_OK;
}

static int make_gates(ufem_net *N, float *const *eb, float *const *g, int nb, void *st)
{
  int l;
  for (l = 0; l < N->L; l++)
    if (ugk_gates((long)nb * N->V[l], N->ch[l], eb[l], N->w + N->offA[l], N->w + N->offB[l], g[l], st))
      return net_fail("gates");
  return AF_OK;
}

/* ebar levels on the host: node level, then 3x3x3 means around 2c */
static void pool_host(const float *f, ugk_grid3 gf, float *c, ugk_grid3 gc)
{
  int z, y, x, dz, dy, dx;
  for (z = 0; z < gc.D; z++)
    for (y = 0; y < gc.H; y++)
      for (x = 0; x < gc.W; x++) {
        double s = 0.0;
        int k = 0;
        for (dz = -1; dz <= 1; dz++)
          for (dy = -1; dy <= 1; dy++)
            for (dx = -1; dx <= 1; dx++) {
              const int a = 2 * z + dz, b = 2 * y + dy, e = 2 * x + dx;
              if (a < 0 || a >= gf.D || b < 0 || b >= gf.H || e < 0 || e >= gf.W) continue;
              s += f[((long)a * gf.H + b) * gf.W + e];
              k++;
            }
        c[((long)z * gc.H + y) * gc.W + x] = (float)(s / k);
      }
}

static void levels_host(const ufem_net *N, const float *e0, float *out)
{
  long off = 0;
  int l;
  if (N->fine) memcpy(out, e0, sizeof(float) * (size_t)N->V[0]);
  else pool_host(e0, N->gf, out, N->g[0]);
  for (l = 1; l < N->L; l++) {
    pool_host(out + off, N->g[l - 1], out + off + N->V[l - 1], N->g[l]);
    off += N->V[l - 1];
  }
}

/* upload nb samples' ebar0 (host [nb][V0]) into the batch level buffers */
static int upload_levels(ufem_net *N, const float *eb0, int nb)
{
  float *lv = (float *)malloc(sizeof(float) * (size_t)N->Vt);
  int b, l, rv = AF_OK;
  if (!lv) return AF_ERR_MEM;
  for (b = 0; b < nb && rv == AF_OK; b++) {
    long off = 0;
    levels_host(N, eb0 + b * N->Vf, lv);
    for (l = 0; l < N->L; l++) {
      if (ugk_h2d(N->eb[l] + b * N->V[l], lv + off, sizeof(float) * (size_t)N->V[l], N->stream) ||
          ugk_sync(N->stream))
        rv = net_fail("levels upload");
      off += N->V[l];
    }
  }
  free(lv);
  return rv;
}

static int upload_input(ufem_net *N, const float *x, int nb)
{
  const long V = N->Vf;
  float *h = (float *)calloc((size_t)(nb * V * 4), sizeof(float));
  long i;
  int c, rv;
  if (!h) return AF_ERR_MEM;
  for (i = 0; i < nb * V; i++)
    for (c = 0; c < 3; c++) h[i * 4 + c] = x[i * 3 + c];
  rv = ugk_h2d(N->x0, h, sizeof(float) * (size_t)(nb * V * 4), N->stream) || ugk_sync(N->stream);
  free(h);
  return rv ? net_fail("upload") : AF_OK;
}

int ufem_net_forward(ufem_net *N, int nb, const float *x, const float *eb0, float *y)
{
  const long V = N->Vf;
  float *h;
  long i;
  int c;
  int b;
  if (nb < 1 || nb > N->nbmax) { ufem_set_error("batch %d out of range", nb); return AF_ERR_VAL; }
  if (upload_input(N, x, nb) || upload_levels(N, eb0, nb) || make_gates(N, N->eb, N->gt, nb, N->stream))
    return AF_ERR;
  /* the inference path (split activations), one sample at a time */
  for (b = 0; b < nb; b++) {
    float *g[NL_MAX];
    int l;
    for (l = 0; l < N->L; l++) g[l] = N->gt[l] + (long)b * N->V[l] * N->ch[l];
    if (ugk_to_split(V, N->x0 + (long)b * V * 4, N->x0s, N->stream)) return net_fail("split input");
    if (infer(N, g, N->y + (long)b * V * 4, NULL, N->stream)) return AF_ERR;
  }
  h = (float *)malloc(sizeof(float) * (size_t)(nb * V * 4));
  if (!h) return AF_ERR_MEM;
  if (ugk_d2h(h, N->y, sizeof(float) * (size_t)(nb * V * 4), N->stream)) { free(h); return net_fail("download"); }
  for (i = 0; i < nb * V; i++)
    for (c = 0; c < 3; c++) y[i * 3 + c] = h[i * 4 + c];
  free(h);
  return AF_OK;
}

/* gradient of <dy, N(x)> with respect to the parameters (packed like the
 * weights), for nb samples */
int ufem_net_backward(ufem_net *N, int nb, const float *x, const float *eb0, const float *dy, float *grad)
{
  const long V = N->Vf;
  float *h;
  long i;
  int c, rv;
  if (nb < 1 || nb > N->nbmax) { ufem_set_error("batch %d out of range", nb); return AF_ERR_VAL; }
  h = (float *)calloc((size_t)(nb * V * 4), sizeof(float));
  if (!h) return AF_ERR_MEM;
  /* the training forward (f32 activations kept for the backward) */
  rv = (upload_input(N, x, nb) || upload_levels(N, eb0, nb) || make_gates(N, N->eb, N->gt, nb, N->stream) ||
        net_forward(N, N->x0, N->gt, N->y, nb, NULL, N->stream)) ? AF_ERR : AF_OK;
  if (rv == AF_OK) {
    for (i = 0; i < nb * V; i++)
      for (c = 0; c < 3; c++) h[i * 4 + c] = dy[i * 3 + c];
    if (ugk_h2d(N->dy, h, sizeof(float) * (size_t)(nb * V * 4), N->stream) || ugk_sync(N->stream))
      rv = net_fail("upload");
  }
  if (rv == AF_OK) rv = net_backward(N, N->x0, N->dy, nb);
  if (rv == AF_OK && ugk_d2h(grad, N->gr, sizeof(float) * N->P, N->stream)) rv = net_fail("download");
  free(h);
  return rv;
}

/* ---------------- binding, training pairs, training ---------------------- */
int ufem_net_bind(ufem_net *N, ufem_sys *s, const double *E)
{
  const ugk_grid3 g = N->gf;
  const long V = N->Vf, nx = g.W - 1, ny = g.H - 1, nz = g.D - 1;
  const long Vt = N->Vt;
  double *acc;
  int *cnt;
  float *e0, *mk;
  long e, i;
  int l, rv = AF_OK;
  if (s->n != 3 * V || s->n_elems != nx * ny * nz) {
    ufem_set_error("system (%ld DOF, %ld bricks) does not match the %dx%dx%d node grid", s->n, s->n_elems,
                   g.W, g.H, g.D);
    return AF_ERR_VAL;
  }
  acc = (double *)calloc((size_t)V, sizeof(double));
  cnt = (int *)calloc((size_t)V, sizeof(int));
  e0 = (float *)malloc(sizeof(float) * (size_t)V);
  mk = (float *)malloc(sizeof(float) * (size_t)s->n);
  if (!N->ebh) N->ebh = (float *)malloc(sizeof(float) * (size_t)Vt);
  if (!acc || !cnt || !e0 || !mk || !N->ebh) { rv = AF_ERR_MEM; goto out; }
  /* brick e = i + nx (j + ny k) touches nodes (i+a) + W ((j+b) + H (k+c)) */
  for (e = 0; e < nx * ny * nz; e++) {
    const long ii = e % nx, jj = (e / nx) % ny, kk = e / (nx * ny);
    const double le = log(E[e] / E_REF);
    int a, b, c;
    for (c = 0; c < 2; c++)
      for (b = 0; b < 2; b++)
        for (a = 0; a < 2; a++) {
          const long p = (ii + a) + g.W * ((jj + b) + g.H * (kk + c));
          acc[p] += le;
          cnt[p]++;
        }
  }
  for (i = 0; i < V; i++) e0[i] = (float)(cnt[i] ? acc[i] / cnt[i] : 0.0);
  levels_host(N, e0, N->ebh);
  if (!N->mask) N->mask = (float *)ugk_malloc(sizeof(float) * (size_t)s->n);
  if (!N->mask) { rv = net_fail("mask"); goto out; }
  for (i = 0; i < s->n; i++) mk[i] = 1.f;
  for (i = 0; i < s->n_fixed; i++) mk[s->fixed_host[i]] = 0.f;
  if (ugk_h2d(N->mask, mk, sizeof(float) * (size_t)s->n, N->stream) ||
      ugk_h2d(N->ebi, N->ebh, sizeof(float) * (size_t)Vt, N->stream) || ugk_sync(N->stream)) {
    rv = net_fail("upload");
    goto out;
  }
  N->sys = s;
  (void)l;
out:
  free(acc); free(cnt); free(e0); free(mk);
  return rv;
}

int ufem_net_reserve(ufem_net *N, long cap)
{
  const long V = N->Vf, n = 3 * V;
  float *pv, *prr, *pee;
  double *per;
  int *pf;
  if (cap <= N->cap) return AF_OK;
  pv = (float *)ugk_malloc(sizeof(float) * (size_t)(V * 4) * (size_t)cap);
  prr = (float *)ugk_malloc(sizeof(float) * (size_t)n * (size_t)cap);
  pee = (float *)ugk_malloc(sizeof(float) * (size_t)n * (size_t)cap);
  per = (double *)calloc((size_t)cap, sizeof(double));
  pf = (int *)calloc((size_t)cap, sizeof(int));
  if (!pv || !prr || !pee || !per || !pf) {
    ugk_free(pv); ugk_free(prr); ugk_free(pee); free(per); free(pf);
    return net_fail("pair store");
  }
  if (N->count && (ugk_d2d(pv, N->pv, sizeof(float) * (size_t)(V * 4 * N->count), N->stream) ||
                   ugk_d2d(prr, N->prr, sizeof(float) * (size_t)(n * N->count), N->stream) ||
                   ugk_d2d(pee, N->pee, sizeof(float) * (size_t)(n * N->count), N->stream) ||
                   ugk_sync(N->stream))) {
    ugk_free(pv); ugk_free(prr); ugk_free(pee); free(per); free(pf);
    return net_fail("pair store copy");
  }
  if (N->count) {
    memcpy(per, N->per, sizeof(double) * (size_t)N->count);
    memcpy(pf, N->pfield, sizeof(int) * (size_t)N->count);
  }
  ugk_free(N->pv); ugk_free(N->prr); ugk_free(N->pee); free(N->per); free(N->pfield);
  N->pv = pv; N->prr = prr; N->pee = pee; N->per = per; N->pfield = pf;
  N->cap = cap;
  return AF_OK;
}

long ufem_net_npairs(const ufem_net *N) { return N->count; }

/* chunk hook of the PCG driver: snapshot (x, r) every snap_every iterations */
static int snap_hook(void *ctx, long it, void *stream)
{
  ufem_net *N = (ufem_net *)ctx;
  ufem_sys *s = N->sys;
  const long j = it / N->snap_every;
  const size_t b = sizeof(double) * (size_t)s->n;
  if (it % N->snap_every || j >= N->snap_cap) return 0;
  if (ugk_d2d(N->snap + 2 * j * s->n, s->d_u, b, stream) || ugk_d2d(N->snap + (2 * j + 1) * s->n, s->w.r, b, stream))
    return -1;
  if (j + 1 > N->snap_n) N->snap_n = j + 1;
  return 0;
}

/* keep the bound system's current material field (FP32 K, dinv, ebar) */
static int add_field(ufem_net *N)
{
  ufem_sys *s = N->sys;
  const long Vt = N->Vt;
  const size_t nv = (size_t)s->K.total_slots * 64 * 9;
  ugk_mat tmp;
  int f = N->nfield;
  if (f == N->capf) {
    int cap = N->capf ? 2 * N->capf : 16;
    float **a = (float **)realloc(N->fvals, sizeof(float *) * (size_t)cap);
    float **b = a ? (float **)realloc(N->fdinv, sizeof(float *) * (size_t)cap) : NULL;
    float **c = b ? (float **)realloc(N->feb, sizeof(float *) * (size_t)cap) : NULL;
    if (a) N->fvals = a;
    if (b) N->fdinv = b;
    if (c) N->feb = c;
    if (!a || !b || !c) return AF_ERR_MEM;
    N->capf = cap;
  }
  N->fvals[f] = (float *)ugk_malloc(sizeof(float) * nv);
  N->fdinv[f] = (float *)ugk_malloc(sizeof(float) * (size_t)s->n);
  N->feb[f] = (float *)ugk_malloc(sizeof(float) * (size_t)Vt);
  if (!N->fvals[f] || !N->fdinv[f] || !N->feb[f]) {
    ugk_free(N->fvals[f]); ugk_free(N->fdinv[f]); ugk_free(N->feb[f]);
    return net_fail("field store");
  }
  tmp = s->K;
  tmp.d_vals32 = N->fvals[f];
  tmp.d_corr32 = NULL;    /* the training loss uses the plain FP32 copy */
  tmp.d_rem16 = NULL;
  if (ugk_vals_to_f32(&tmp, N->stream) || ugk_to_f32(s->n, s->w.dinv, N->fdinv[f], N->stream) ||
      ugk_h2d(N->feb[f], N->ebh, sizeof(float) * (size_t)Vt, N->stream) || ugk_sync(N->stream))
    return net_fail("field store");
  N->nfield++;
  return AF_OK;
}

long ufem_net_collect(ufem_net *N, int pairs, double eps, int every, ufem_solve_info *info)
{
  ufem_sys *s = N->sys;
  ufem_solve_opts o;
  ufem_solve_info tmp;
  long added = 0, jmax, last = -1, k;
  double er0 = 0.0;
  int rv, f;
  if (!s) { ufem_set_error("bind the network to a system first"); return AF_ERR; }
  if (pairs < 1 || every < 2 || every % 2) { ufem_set_error("pairs >= 1, even every >= 2"); return AF_ERR_VAL; }
  if (!info) info = &tmp;
  if (!N->snap) {
    long cap = (long)(4e9 / (16.0 * (double)s->n));
    N->snap_cap = cap < 16 ? 16 : (cap > 512 ? 512 : cap);
    N->snap = (double *)ugk_malloc(sizeof(double) * 2 * (size_t)s->n * (size_t)N->snap_cap);
    if (!N->snap) return net_fail("snapshot ring");
  }
  N->snap_every = every;
  N->snap_n = 0;
  memset(&o, 0, sizeof(o));
  o.solver = UFEM_SOLVER_PCG;
  o.rule = UFEM_RULE_RELATIVE;
  o.eps = eps;
  o.exec = UFEM_EXEC_STREAMED;
  o.chunk = every;
  s->w.hook = snap_hook;
  s->w.hook_ctx = N;
  rv = ufem_sys_solve(s, &o, info);
  s->w.hook = NULL;
  s->w.hook_ctx = NULL;
  if (rv != AF_OK) return rv;
  if (add_field(N)) return AF_ERR;
  f = N->nfield - 1;
  jmax = info->iters / every;
  if (jmax > N->snap_n - 1) jmax = N->snap_n - 1;
  for (k = 0; k < pairs && N->count < N->cap; k++) {
    const long j = pairs == 1 ? 0 : (long)floor((double)k * (double)jmax / (double)(pairs - 1) + 0.5);
    const double *x = N->snap + 2 * j * s->n, *r = N->snap + (2 * j + 1) * s->n;
    double a2 = 0.0, er = 0.0, sc;
    int G, i;
    if (j == last) continue;
    last = j;
    G = ugk_pair_stats(s->n, x, r, s->d_u, s->w.dinv, N->dpart, N->stream);
    if (G < 0 || ugk_d2h(N->hpart, N->dpart, sizeof(double) * 2 * (size_t)G, N->stream))
      return net_fail("pair statistics");
    for (i = 0; i < G; i++) {
      a2 += N->hpart[2 * i];
      er += N->hpart[2 * i + 1];
    }
    if (k == 0) er0 = er;
    /* skip converged snapshots, whose error is at the solve's own accuracy */
    if (!(a2 > 0.0) || !(er > 1e-14 * er0)) continue;
    sc = sqrt(a2 / (double)s->n);
    if (ugk_pair_write(s->n, x, r, s->d_u, s->w.dinv, N->mask, sc, N->pv + N->count * N->Vf * 4,
                       N->prr + N->count * s->n, N->pee + N->count * s->n, N->stream))
      return net_fail("pair write");
    N->per[N->count] = er / (sc * sc);
    N->pfield[N->count] = f;
    N->count++;
    added++;
  }
  if (ugk_sync(N->stream)) return net_fail("sync");
  return added;
}

/* energy loss and its output gradient for batch slot b holding pair q */
static int pair_loss(ufem_net *N, int b, long q, int nb, double *part_out)
{
  ufem_sys *s = N->sys;
  const long n = s->n;
  const int f = N->pfield[q];
  ugk_mat Kf = s->K;
  int G;
  Kf.d_vals32 = N->fvals[f];
  Kf.d_corr32 = NULL;
  Kf.d_rem16 = NULL;
  if (ugk_el_z(n, N->prr + q * n, N->fdinv[f], N->mask, N->y + b * N->Vf * 4, N->zz, N->stream) ||
      ugk_spmv_f32(&Kf, N->zz, N->kz, N->stream))
    return net_fail("loss");
  G = ugk_el_grad(n, N->prr + q * n, N->pee + q * n, N->zz, N->kz, N->fdinv[f], N->mask,
                  (float)(-2.0 / (N->per[q] * nb)), N->dy + b * N->Vf * 4, part_out, N->stream);
  if (G < 0) return net_fail("loss");
  return G;
}

/* load pairs idx[0..nb) into the batch: inputs, ebar levels, gates */
static int load_batch(ufem_net *N, const long *idx, int nb)
{
  int b, l;
  for (b = 0; b < nb; b++) {
    const long q = idx[b];
    long off = 0;
    if (ugk_d2d(N->x0 + b * N->Vf * 4, N->pv + q * N->Vf * 4, sizeof(float) * (size_t)(N->Vf * 4),
                N->stream))
      return net_fail("batch copy");
    for (l = 0; l < N->L; l++) {
      if (ugk_d2d(N->eb[l] + b * N->V[l], N->feb[N->pfield[q]] + off, sizeof(float) * (size_t)N->V[l], N->stream))
        return net_fail("batch copy");
      off += N->V[l];
    }
  }
  return make_gates(N, N->eb, N->gt, nb, N->stream);
}

/* losses of pairs [i0, i0 + cnt) under the current weights (no update) */
int ufem_net_eval(ufem_net *N, long i0, long cnt, double *loss)
{
  long k;
  int G, i;
  if (!N->sys || i0 < 0 || i0 + cnt > N->count) { ufem_set_error("pair range"); return AF_ERR_VAL; }
  for (k = 0; k < cnt; k++) {
    long q = i0 + k;
    double sum = 0.0;
    if (load_batch(N, &q, 1) || net_forward(N, N->x0, N->gt, N->y, 1, NULL, N->stream)) return AF_ERR;
    G = pair_loss(N, 0, q, 1, N->dpart);
    if (G < 0 || ugk_d2h(N->hpart, N->dpart, sizeof(double) * (size_t)G, N->stream)) return AF_ERR;
    for (i = 0; i < G; i++) sum += N->hpart[i];
    loss[k] = sum / N->per[q];
  }
  return AF_OK;
}

int ufem_net_train(ufem_net *N, long steps, int batch, double lr, unsigned long seed, double *loss, double *ms)
{
  long *perm = NULL, pos, st, idx[64];
  unsigned long long rng = (unsigned long long)seed * 0x9E3779B97F4A7C15ULL + 1;
  int rv = AF_OK;
  if (!N->sys) { ufem_set_error("bind the network to a system first"); return AF_ERR; }
  if (batch < 1 || batch > N->nbmax || batch > 64) { ufem_set_error("batch %d out of range", batch); return AF_ERR_VAL; }
  if (N->count < batch) { ufem_set_error("%ld training pairs, batch %d", N->count, batch); return AF_ERR_VAL; }
  perm = (long *)malloc(sizeof(long) * (size_t)N->count);
  if (!perm) return AF_ERR_MEM;
  pos = N->count;
  if (ugk_event_record(N->ev[0], N->stream)) { free(perm); return net_fail("event"); }
  for (st = 0; st < steps; st++) {
    /* cosine decay from lr to lr / 10 */
    const double f = 0.55 + 0.45 * cos(3.141592653589793 * (double)st / (double)(steps > 1 ? steps - 1 : 1));
    int k, G[64];
    for (k = 0; k < batch; k++) {
      if (pos >= N->count) {   /* new epoch: Fisher-Yates shuffle */
        long i;
        for (i = 0; i < N->count; i++) perm[i] = i;
        for (i = N->count - 1; i > 0; i--) {
          long jj, tq;
          rng = mix64(rng);
          jj = (long)(rng % (unsigned long long)(i + 1));
          tq = perm[i]; perm[i] = perm[jj]; perm[jj] = tq;
        }
        pos = 0;
      }
      idx[k] = perm[pos++];
    }
    if (load_batch(N, idx, batch) || net_forward(N, N->x0, N->gt, N->y, batch, NULL, N->stream)) {
      rv = AF_ERR;
      goto out;
    }
    for (k = 0; k < batch; k++) {
      G[k] = pair_loss(N, k, idx[k], batch, N->dpart + 1024L * k);
      if (G[k] < 0) { rv = AF_ERR; goto out; }
    }
    if (net_backward(N, N->x0, N->dy, batch) || adam_all(N, (float)(lr * f))) { rv = AF_ERR; goto out; }
    if (loss) {
      double sum = 0.0;
      if (ugk_d2h(N->hpart, N->dpart, sizeof(double) * 1024 * (size_t)batch, N->stream)) {
        rv = net_fail("loss");
        goto out;
      }
      for (k = 0; k < batch; k++) {
        double sk = 0.0;
        int i;
        for (i = 0; i < G[k]; i++) sk += N->hpart[1024L * k + i];
        sum += sk / N->per[idx[k]];
      }
      loss[st] = sum / batch;
    }
  }
  if (ugk_event_record(N->ev[1], N->stream) || ugk_sync(N->stream)) { rv = net_fail("sync"); goto out; }
  if (ms) *ms = ugk_event_ms(N->ev[0], N->ev[1]);
out:
  free(perm);
  return rv;
}

/* ---------------- inference inside flexible CG (ufem_solve.c) ------------- */
long ufem_net_bound_n(const ufem_net *N) { return N->sys ? N->sys->n : 0; }
unsigned short *ufem_net_input(ufem_net *N) { return N->x0s; }
const float *ufem_net_mask(const ufem_net *N) { return N->mask; }
const float *ufem_net_output(const ufem_net *N) { return N->yi; }

/* gates of the bound field, before a solve (the weights may have changed) */
int ufem_net_prepare(ufem_net *N, void *stream)
{
  float *eb[NL_MAX];
  long off = 0;
  int l;
  for (l = 0; l < N->L; l++) {
    eb[l] = N->ebi + off;
    off += N->V[l];
  }
  return make_gates(N, eb, N->gi, 1, stream);
}

int ufem_net_infer(ufem_net *N, const int *skip, void *stream)
{
  return infer(N, N->gi, N->yi, skip, stream);
}

/* the last stage, y = C0^T t0, as a convolution descriptor for a fused
 * consumer (the coarse-first form only: there C0^T is a stride-2
 * transpose); 0 when the form has no such stage */
int ufem_net_last_stage(const ufem_net *N, ugk_conv_desc *d)
{
  if (N->fine || (N->ch[0] != 16 && N->ch[0] != 32)) return 0;
  memset(d, 0, sizeof(*d));
  d->x = N->t[0]; d->ldx = N->ch[0]; d->cin = N->ch[0]; d->x_f32 = UGK_ACT_F32;
  d->w = N->wt[0]; d->cip = r16(N->ch[0]); d->cop = r16(N->ci[0]);
  d->y = N->yi; d->ldy = 4; d->cout = 3; d->y_f32 = 1;
  d->mode = UGK_CONV_T2;
  d->nb = 1;
  d->gi = N->g[0];
  d->go = N->gf;
  return 1;
}

/* all inference stages but the last */
int ufem_net_infer_head(ufem_net *N, const int *skip, void *stream)
{
  int k;
  for (k = 0; k < infer_stages(N) - 1; k++)
    if (infer_stage(N, k, N->gi, N->yi, skip, stream)) return net_fail("network inference");
  return AF_OK;
}

/* ---- FP64 inference --------------------------------------------------------
 * The solver applies the network with FP64 activations and sums (the same
 * operator: bf16 weights widened exactly, f32 gates): flexible CG with one
 * kept direction needs one linear map from iteration to iteration, and the
 * matrix cores' f32 rounding noise costs it iterations.  The coarse-first
 * form with 16 or 32 channels per level. */
int ufem_net_f64_ok(const ufem_net *N)
{
  int l;
  if (N->fine || N->ci[0] != 3) return 0;
  for (l = 0; l < N->L; l++)
    if (N->ch[l] != 16 && N->ch[l] != 32) return 0;
  return 1;
}

double *ufem_net_input64(ufem_net *N)
{
  if (!N->v64) N->v64 = (double *)ugk_malloc(sizeof(double) * 3 * (size_t)N->Vf);
  return N->v64;
}

/* buffers and the FP64 copies of the current weights */
int ufem_net_prepare64(ufem_net *N, void *stream)
{
  int l, bad = 0;
  if (!ufem_net_f64_ok(N)) {
    ufem_set_error("FP64 inference: the coarse-first form with 16 or 32 channels per level only");
    return AF_ERR_VAL;
  }
  if (!ufem_net_input64(N)) return net_fail("FP64 input");
  for (l = 0; l < N->L; l++) {
    if (!N->h64[l]) N->h64[l] = (double *)dalloc(sizeof(double) * (size_t)N->V[l] * N->ch[l], &bad);
    if (!N->t64[l]) N->t64[l] = (double *)dalloc(sizeof(double) * (size_t)N->V[l] * N->ch[l], &bad);
    if (!N->fd64[l]) N->fd64[l] = (double *)dalloc(sizeof(double) * (size_t)ugk_n64_frag_len(N->ci[l], N->ch[l]), &bad);
    if (l && !N->fu64[l]) N->fu64[l] = (double *)dalloc(sizeof(double) * (size_t)ugk_n64_frag_len(N->ch[l], N->ci[l]), &bad);
  }
  if (!N->wu64) N->wu64 = (double *)dalloc(sizeof(double) * 27 * (size_t)N->ch[0] * N->ci[0], &bad);
  if (bad) return net_fail("FP64 activations");
  for (l = 0; l < N->L; l++) {
    const int cop = r16(N->ch[l]), cip = r16(N->ci[l]);
    if (ugk_w64_frag(N->wf[l], cop, cip, N->ch[l], N->ci[l], 0, N->fd64[l], stream) ||
        (l && ugk_w64_frag(N->wf[l], cop, cip, N->ch[l], N->ci[l], 1, N->fu64[l], stream)))
      return net_fail("FP64 weights");
  }
  if (ugk_w64(N->wf[0], r16(N->ch[0]), r16(N->ci[0]), N->ch[0], N->ci[0], N->wu64, stream)) return net_fail("FP64 weights");
  return AF_OK;
}

/* all stages but the last from the input v64 (mask sqrt(dinv) r, [V][3]):
 * h0 = C0 v, h_l = C_l h_{l-1}, t_{L-1} = g_{L-1} h_{L-1},
 * t_l = g_l h_l + C_{l+1}^T t_{l+1} */
int ufem_net_infer64_head(ufem_net *N, const int *skip, void *stream)
{
  const int L = N->L;
  int l;
  for (l = 0; l < L; l++) {
    const double *x = l ? N->h64[l - 1] : N->v64;
    const int cin = l ? N->ch[l - 1] : 3;
    const ugk_grid3 gi = l ? N->g[l - 1] : N->gf;
    if (ugk_n64_down(x, cin, cin, gi, N->g[l], N->fd64[l], N->ch[l], l == L - 1 ? N->gi[l] : NULL,
                     l == L - 1 ? N->t64[l] : N->h64[l], skip, stream))
      return net_fail("FP64 inference");
  }
  for (l = L - 2; l >= 0; l--)
    if (ugk_n64_up(N->t64[l + 1], N->ch[l + 1], N->g[l + 1], N->g[l], N->fu64[l + 1], N->ch[l], N->gi[l], N->h64[l],
                   N->t64[l], skip, stream))
      return net_fail("FP64 inference");
  return AF_OK;
}

/* the last stage for the fused combine: y = C0^T t0 from the FP64 t0 */
int ufem_net_last64(const ufem_net *N, ugk_conv_desc *d)
{
  if (!ufem_net_last_stage(N, d)) return 0;
  d->x = N->t64[0];
  d->x_f32 = UGK_ACT_F64;
  return 1;
}

/* y = N(v) with FP64 inference for the bound field: v, y [D][H][W][3] (host) */
int ufem_net_apply64(ufem_net *N, const double *v, double *y)
{
  const long n = 3 * N->Vf;
  int rv;
  if (!N->sys) { ufem_set_error("ufem_net_apply64: no bound system"); return AF_ERR_VAL; }
  if ((rv = ufem_net_prepare64(N, N->stream)) != AF_OK) return rv;
  if (!N->y64) N->y64 = (double *)ugk_malloc(sizeof(double) * (size_t)n);
  if (!N->y64) return net_fail("FP64 output");
  if (ugk_h2d(N->v64, v, sizeof(double) * (size_t)n, N->stream) || ufem_net_prepare(N, N->stream) ||
      ufem_net_infer64_head(N, NULL, N->stream) ||
      ugk_n64_out(N->t64[0], N->ch[0], N->g[0], N->gf, N->wu64, N->y64, N->stream) ||
      ugk_d2h(y, N->y64, sizeof(double) * (size_t)n, N->stream) || ugk_sync(N->stream))
    return net_fail("FP64 inference");
  return AF_OK;
}

/* pair i as stored: v [V][4], rr [n], ee [n], e.r and its field index */
int ufem_net_get_pair(ufem_net *N, long i, float *v, float *rr, float *ee, double *er, int *field)
{
  const long n = 3 * N->Vf;
  if (i < 0 || i >= N->count) { ufem_set_error("pair %ld out of range", i); return AF_ERR_VAL; }
  if (ugk_d2h(v, N->pv + i * N->Vf * 4, sizeof(float) * (size_t)(N->Vf * 4), N->stream) ||
      ugk_d2h(rr, N->prr + i * n, sizeof(float) * (size_t)n, N->stream) ||
      ugk_d2h(ee, N->pee + i * n, sizeof(float) * (size_t)n, N->stream))
    return net_fail("pair download");
  *er = N->per[i];
  *field = N->pfield[i];
  return AF_OK;
}

/* ---------------- single kernels, for tests ------------------------------- */
static unsigned short h_f2bf(float f)
{
  unsigned u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

/* y = conv(x) in mode (0 stride 1, 1 stride 2, 2 stride-2 transpose) with
 * weights w [27][cout][cin] (f32 out), x bf16-rounded or f32 (x_f32); when
 * dy is given, also dw [27][cout][cin] = the weight gradient (modes 0, 1) */
int ufem_net_probe_conv(int mode, int flip, int x_f32, int nb, const int *gi, const int *go, int cin, int cout,
                        const float *x, const float *w, float *y, const float *dy, float *dw)
{
  ugk_grid3 Gi, Go;
  long Vi, Vo, i;
  const int cip = r16(cin), cop = r16(cout), ldx = (cin + 3) / 4 * 4, ldy = (cout + 3) / 4 * 4;
  const size_t xb = (x_f32 ? sizeof(float) : sizeof(unsigned short)) * (size_t)ldx;
  void *hx = NULL, *dx_ = NULL, *dxs = NULL;
  unsigned short *hw = NULL, *dw_ = NULL;
  float *dy_ = NULL, *part = NULL, *dgr = NULL, *hy = NULL, *hd = NULL, *ddy = NULL;
  void *st = NULL;
  int rv = AF_ERR, c, t, o;
  Gi.D = gi[0]; Gi.H = gi[1]; Gi.W = gi[2];
  Go.D = go[0]; Go.H = go[1]; Go.W = go[2];
  if (Gi.D < 1 || Gi.H < 1 || Gi.W < 1 || Go.D < 1 || Go.H < 1 || Go.W < 1 || nb < 1 || cin < 1 || cout < 1 ||
      cip > 96 || cop > 96) {
    ufem_set_error("probe shape");
    return AF_ERR_VAL;
  }
  if (ugk_init()) return net_fail("device");
  Vi = (long)Gi.D * Gi.H * Gi.W;
  Vo = (long)Go.D * Go.H * Go.W;
  st = ugk_stream_create();
  hx = calloc((size_t)(nb * Vi), xb);
  hw = (unsigned short *)calloc((size_t)27 * cop * cip, sizeof(unsigned short));
  hy = (float *)malloc(sizeof(float) * (size_t)(nb * Vo * ldy));
  if (!st || !hx || !hw || !hy) goto out;
  for (i = 0; i < nb * Vi; i++)
    for (c = 0; c < cin; c++) {
      if (x_f32) ((float *)hx)[i * ldx + c] = x[i * cin + c];
      else ((unsigned short *)hx)[i * ldx + c] = h_f2bf(x[i * cin + c]);
    }
  for (t = 0; t < 27; t++)
    for (o = 0; o < cout; o++)
      for (c = 0; c < cin; c++) hw[((long)t * cop + o) * cip + c] = h_f2bf(w[((long)t * cout + o) * cin + c]);
  dx_ = ugk_malloc(xb * (size_t)(nb * Vi));
  dw_ = (unsigned short *)ugk_malloc(sizeof(unsigned short) * (size_t)27 * cop * cip);
  dy_ = (float *)ugk_malloc(sizeof(float) * (size_t)(nb * Vo * ldy));
  if (!dx_ || !dw_ || !dy_) goto out;
  if (ugk_h2d(dx_, hx, xb * (size_t)(nb * Vi), st) || ugk_h2d(dw_, hw, sizeof(unsigned short) * 27 * cop * cip, st))
    goto out;
  if (x_f32 == UGK_ACT_SPLIT) {     /* the f32 input, stored split */
    dxs = ugk_malloc(3 * sizeof(unsigned short) * (size_t)(nb * Vi * ldx));
    if (!dxs || ugk_to_split(nb * Vi * ldx / 4, (const float *)dx_, (unsigned short *)dxs, st)) goto out;
  }
  {
    ugk_conv_desc d;
    memset(&d, 0, sizeof(d));
    d.x = dxs ? dxs : dx_; d.ldx = ldx; d.cin = ldx; d.x_f32 = x_f32;
    d.w = dw_; d.cip = cip; d.cop = cop;
    d.y = dy_; d.ldy = ldy; d.cout = cout; d.y_f32 = 1;
    d.mode = mode; d.flip = flip; d.nb = nb; d.gi = Gi; d.go = Go;
    if (ugk_conv(&d, st) || ugk_d2h(hy, dy_, sizeof(float) * (size_t)(nb * Vo * ldy), st)) goto out;
  }
  for (i = 0; i < nb * Vo; i++)
    for (c = 0; c < cout; c++) y[i * cout + c] = hy[i * ldy + c];
  if (dy && dw) {
    ugk_dw_desc d;
    int ks;
    hd = (float *)calloc((size_t)(nb * Vo * ldy), sizeof(float));
    ddy = (float *)ugk_malloc(sizeof(float) * (size_t)(nb * Vo * ldy));
    if (!hd || !ddy) goto out;
    for (i = 0; i < nb * Vo; i++)
      for (c = 0; c < cout; c++) hd[i * ldy + c] = dy[i * cout + c];
    if (ugk_h2d(ddy, hd, sizeof(float) * (size_t)(nb * Vo * ldy), st)) goto out;
    memset(&d, 0, sizeof(d));
    d.x = dx_; d.ldx = ldx; d.cin = cin; d.x_f32 = x_f32 ? 1 : 0;
    d.dy = ddy; d.ldd = ldy; d.cout = cout; d.dy_f32 = 1;
    d.cip = cip; d.cop = cop; d.mode = mode; d.nb = nb; d.gi = Gi; d.go = Go;
    ks = ugk_conv_dw_slices(&d);
    part = (float *)ugk_malloc(sizeof(float) * (size_t)ks * 27 * cop * cip);
    dgr = (float *)ugk_malloc(sizeof(float) * (size_t)27 * cout * cin);
    if (!part || !dgr) goto out;
    if (ugk_conv_dw(&d, part, ks, dgr, 0, st) ||
        ugk_d2h(dw, dgr, sizeof(float) * (size_t)27 * cout * cin, st))
      goto out;
  }
  rv = AF_OK;
out:
  if (rv != AF_OK) net_fail("probe");
  free(hx); free(hw); free(hy); free(hd);
  ugk_free(dx_); ugk_free(dxs); ugk_free(dw_); ugk_free(dy_); ugk_free(ddy); ugk_free(part); ugk_free(dgr);
  if (st) ugk_stream_destroy(st);
  return rv;
}

/* average device time of `reps` launches of one convolution on random data
 * (for tuning and for the roofline of the network kernels) */
int ufem_net_time_conv(int mode, int flip, int x_f32, int nb, const int *gi, const int *go, int cin, int cout,
                       int reps, double *us)
{
  ugk_grid3 Gi, Go;
  long Vi, Vo;
  const int cip = r16(cin), cop = r16(cout), ldx = (cin + 3) / 4 * 4, ldy = (cout + 3) / 4 * 4;
  void *x = NULL, *st = NULL, *e0 = NULL, *e1 = NULL;
  unsigned short *w = NULL;
  float *y = NULL;
  ugk_conv_desc d;
  int r, rv = AF_ERR;
  Gi.D = gi[0]; Gi.H = gi[1]; Gi.W = gi[2];
  Go.D = go[0]; Go.H = go[1]; Go.W = go[2];
  if (ugk_init()) return net_fail("device");
  Vi = (long)Gi.D * Gi.H * Gi.W;
  Vo = (long)Go.D * Go.H * Go.W;
  st = ugk_stream_create();
  e0 = ugk_event_create();
  e1 = ugk_event_create();
  x = ugk_malloc((x_f32 == UGK_ACT_SPLIT ? 6 : (x_f32 ? 4 : 2)) * (size_t)(nb * Vi * ldx));
  w = (unsigned short *)ugk_malloc(2 * (size_t)27 * cop * cip);
  y = (float *)ugk_malloc(4 * (size_t)(nb * Vo * ldy));
  if (!st || !e0 || !e1 || !x || !w || !y) goto out;
  memset(&d, 0, sizeof(d));
  d.x = x; d.ldx = ldx; d.cin = ldx; d.x_f32 = x_f32;
  d.w = w; d.cip = cip; d.cop = cop;
  d.y = y; d.ldy = ldy; d.cout = cout; d.y_f32 = 1;
  d.mode = mode; d.flip = flip; d.nb = nb; d.gi = Gi; d.go = Go;
  if (ugk_conv(&d, st) || ugk_sync(st)) goto out;
  if (ugk_event_record(e0, st)) goto out;
  for (r = 0; r < reps; r++)
    if (ugk_conv(&d, st)) goto out;
  if (ugk_event_record(e1, st) || ugk_sync(st)) goto out;
  *us = 1000.0 * ugk_event_ms(e0, e1) / reps;
  rv = AF_OK;
out:
  if (rv != AF_OK) net_fail("conv timing");
  ugk_free(x); ugk_free(w); ugk_free(y);
  if (e0) ugk_event_destroy(e0);
  if (e1) ugk_event_destroy(e1);
  if (st) ugk_stream_destroy(st);
  return rv;
}

/* stage k of FP64 inference: C0 .. C_{L-1} (k < L), C_{L-1}^T .. C1^T, and
 * C0^T on its own (k = 2 L - 1; the solver folds it into its combine) */
static int infer64_stage(ufem_net *N, int k)
{
  const int L = N->L;
  if (k < L) {
    const double *x = k ? N->h64[k - 1] : N->v64;
    const int cin = k ? N->ch[k - 1] : 3;
    return ugk_n64_down(x, cin, cin, k ? N->g[k - 1] : N->gf, N->g[k], N->fd64[k], N->ch[k],
                        k == L - 1 ? N->gi[k] : NULL, k == L - 1 ? N->t64[k] : N->h64[k], NULL, N->stream);
  }
  if (k < 2 * L - 1) {
    const int l = 2 * L - 2 - k;
    return ugk_n64_up(N->t64[l + 1], N->ch[l + 1], N->g[l + 1], N->g[l], N->fu64[l + 1], N->ch[l], N->gi[l],
                      N->h64[l], N->t64[l], NULL, N->stream);
  }
  return ugk_n64_out(N->t64[0], N->ch[0], N->g[0], N->gf, N->wu64, N->y64, N->stream);
}

int ufem_net_time_stages64(ufem_net *N, int reps, double *us)
{
  int k, r, rv;
  if (!N->sys) { ufem_set_error("bind the network to a system first"); return AF_ERR; }
  if ((rv = ufem_net_prepare64(N, N->stream)) != AF_OK) return rv;
  if (!N->y64) N->y64 = (double *)ugk_malloc(sizeof(double) * 3 * (size_t)N->Vf);
  if (!N->y64) return net_fail("FP64 output");
  if (ufem_net_prepare(N, N->stream) || ugk_zero(N->v64, sizeof(double) * 3 * (size_t)N->Vf, N->stream))
    return net_fail("FP64 stages");
  for (k = 0; k < 2 * N->L; k++) {
    if (infer64_stage(N, k)) return net_fail("stage");
    if (ugk_event_record(N->ev[0], N->stream)) return net_fail("event");
    for (r = 0; r < reps; r++)
      if (infer64_stage(N, k)) return net_fail("stage");
    if (ugk_event_record(N->ev[1], N->stream)) return net_fail("event");
    us[k] = 1000.0 * ugk_event_ms(N->ev[0], N->ev[1]) / reps;
  }
  return AF_OK;
}

/* average device time (us) of each inference stage on the bound system's
 * grid (us[0 .. 2 L - 1], stages as in infer_stage), `reps` launches each */
int ufem_net_time_stages(ufem_net *N, int reps, double *us)
{
  int k, r;
  if (!N->sys) { ufem_set_error("bind the network to a system first"); return AF_ERR; }
  if (ufem_net_prepare(N, N->stream)) return AF_ERR;
  for (k = 0; k < infer_stages(N); k++) {
    if (infer_stage(N, k, N->gi, N->yi, NULL, N->stream)) return net_fail("stage");
    if (ugk_event_record(N->ev[0], N->stream)) return net_fail("event");
    for (r = 0; r < reps; r++)
      if (infer_stage(N, k, N->gi, N->yi, NULL, N->stream)) return net_fail("stage");
    if (ugk_event_record(N->ev[1], N->stream)) return net_fail("event");
    us[k] = 1000.0 * ugk_event_ms(N->ev[0], N->ev[1]) / reps;
  }
  return AF_OK;
}
