/*
 * ufem_coarse.c -- the two-level (coarse-grid) correction of a box mesh:
 * set-up on the host (grid checks, the per-element Galerkin tables, masks)
 * and the per-solve build of the coarse matrix on the device.
 *
 * No counterpart in the reference: uFEM's solvers precondition with the
 * diagonal or SSOR only (fem/fem_eqs.c).  The correction
 *   z_c = fm P A^-1 P^T fm r,   A = P^T K P,
 * with P trilinear from a coarse node grid (cells of cx x cy x cz bricks),
 * removes the long-wavelength error that the diagonal and the learned
 * network leave, so that CG needs a number of iterations that no longer
 * grows with the mesh size.
 *
 * A is assembled cell by cell.  For identical box elements the stiffness
 * is linear in the Lame constants, K_e = lambda_e K^l + mu_e K^m, so a
 * cell's Galerkin block is sum_p lambda_p G_p^l + mu_p G_p^m with
 * G_p = T_p^T K^(l|m) T_p for its element p and T_p the trilinear weights
 * of the element's nodes in the cell.  K^l and K^m come from the device's
 * own element stiffness kernel (2x2x2 Gauss points, exact for boxes).
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ufem_fem_math.h"
#include "ufem_internal.h"

static int gpu_fail(const char *what)
{
  ufem_set_error("%s: %s", what, ugk_error());
  return AF_ERR;
}

void ufem_coarse_free(ugk_coarse *C)
{
  if (!C) return;
  ugk_free(C->d_G); ugk_free(C->d_KC); ugk_free(C->d_A); ugk_free(C->d_piv); ugk_free(C->d_R);
  ugk_free(C->d_C); ugk_free(C->d_cpart); ugk_free(C->d_g); ugk_free(C->d_y); ugk_free(C->d_fm);
  ugk_free(C->d_cm); ugk_free(C->d_zc);
  free(C);
}

/* brick node order (fem/fem_e009.c:36-69): offsets of node m */
static const int OFF[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                              {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

/* the node grid and element numbering of a box mesh: coordinates
 * x0 + i hx, ..., elements ex + nx (ey + ny ez) with the brick's nodes */
static int check_box(ufem_sys *s, int W, int H, int D, double h[3])
{
  const long nn = s->n_nodes, ne = s->n_elems;
  const long nx = W - 1, ny = H - 1, nz = D - 1;
  double *x = (double *)malloc(sizeof(double) * (size_t)nn * 3);
  int *en = (int *)malloc(sizeof(int) * (size_t)ne * 8);
  long i, e;
  int rv = AF_OK;
  if (!x || !en) { free(x); free(en); return AF_ERR_MEM; }
  if (ugk_d2h(x, s->d_x, sizeof(double) * (size_t)nn, s->stream) ||
      ugk_d2h(x + nn, s->d_y, sizeof(double) * (size_t)nn, s->stream) ||
      ugk_d2h(x + 2 * nn, s->d_z, sizeof(double) * (size_t)nn, s->stream) ||
      ugk_d2h(en, s->d_enodes, sizeof(int) * (size_t)ne * 8, s->stream)) {
    free(x); free(en);
    return gpu_fail("coarse: mesh read-back");
  }
  h[0] = (x[nx] - x[0]) / (double)nx;
  h[1] = (x[nn + (long)W * ny] - x[nn]) / (double)ny;
  h[2] = (x[2 * nn + (long)W * H * nz] - x[2 * nn]) / (double)nz;
  for (i = 0; i < nn && rv == AF_OK; i++) {
    const long a = i % W, b = (i / W) % H, c = i / ((long)W * H);
    const double ex = x[0] + a * h[0], ey = x[nn] + b * h[1], ez = x[2 * nn] + c * h[2];
    const double tol = 1e-9 * (fabs(h[0]) + fabs(h[1]) + fabs(h[2]));
    if (fabs(x[i] - ex) > tol || fabs(x[nn + i] - ey) > tol || fabs(x[2 * nn + i] - ez) > tol) rv = AF_ERR_VAL;
  }
  for (e = 0; e < ne && rv == AF_OK; e++) {
    const long a = e % nx, b = (e / nx) % ny, c = e / (nx * ny);
    int m;
    for (m = 0; m < 8; m++)
      if (en[8 * e + m] != (a + OFF[m][0]) + (long)W * ((b + OFF[m][1]) + (long)H * (c + OFF[m][2]))) rv = AF_ERR_VAL;
  }
  if (rv != AF_OK) ufem_set_error("coarse: the mesh is not a regular box grid of W x H x D nodes");
  if (rv == AF_OK && (h[0] <= 0 || h[1] <= 0 || h[2] <= 0)) {
    ufem_set_error("coarse: the grid must grow along x, y and z");
    rv = AF_ERR_VAL;
  }
  free(x);
  free(en);
  return rv;
}

/* G_p^l, G_p^m for every element position p of a cell */
static int galerkin_tables(int cx, int cy, int cz, const double h[3], double *G)
{
  double xyz[2 * 24], E[2] = {2.0, 1.0}, nu[2] = {0.0, 0.25}, Ke[2 * 576], Kl[576], Km[576];
  int n, a, k, p;
  for (k = 0; k < 2; k++)
    for (n = 0; n < 8; n++)
      for (a = 0; a < 3; a++) xyz[24 * k + 3 * n + a] = OFF[n][a] * h[a];
  if (ufem_gpu_element_stiffness(2, xyz, E, nu, 2, Ke) != AF_OK) return AF_ERR;
  /* E = 2, nu = 0: lambda 0, mu 1.  E = 1, nu = 1/4: lambda = mu = 2/5 */
  for (k = 0; k < 576; k++) {
    Km[k] = Ke[k];
    Kl[k] = (Ke[576 + k] - 0.4 * Km[k]) / 0.4;
  }
  for (p = 0; p < cx * cy * cz; p++) {
    const int pi = p % cx, pj = (p / cx) % cy, pk = p / (cx * cy);
    double T[8][8];   /* T[n][m]: weight of cell corner m at element node n */
    int m, q, mm, b;
    for (n = 0; n < 8; n++)
      for (m = 0; m < 8; m++) {
        const double tx = (double)(pi + OFF[n][0]) / cx, ty = (double)(pj + OFF[n][1]) / cy;
        const double tz = (double)(pk + OFF[n][2]) / cz;
        T[n][m] = (OFF[m][0] ? tx : 1.0 - tx) * (OFF[m][1] ? ty : 1.0 - ty) * (OFF[m][2] ? tz : 1.0 - tz);
      }
    for (q = 0; q < 2; q++) {
      const double *Kq = q ? Km : Kl;
      double *Gp = G + ((size_t)2 * p + q) * 576;
      for (m = 0; m < 8; m++)
        for (a = 0; a < 3; a++)
          for (mm = 0; mm < 8; mm++)
            for (b = 0; b < 3; b++) {
              double acc = 0.0;
              int n2;
              for (n = 0; n < 8; n++) {
                double t = 0.0;
                if (T[n][m] == 0.0) continue;
                for (n2 = 0; n2 < 8; n2++) t += Kq[(3 * n + a) * 24 + 3 * n2 + b] * T[n2][mm];
                acc += T[n][m] * t;
              }
              Gp[(3 * m + a) * 24 + 3 * mm + b] = acc;
            }
    }
  }
  return AF_OK;
}

int ufem_sys_set_coarse(ufem_sys *s, int W, int H, int D, int cx, int cy, int cz)
{
  ugk_coarse *C = NULL;
  double h[3], *G = NULL, *fm = NULL, *cm = NULL;
  long i, n = s->n, nn;
  int rv;
  if (s->w.cc) { ufem_coarse_free(s->w.cc); s->w.cc = NULL; }
  if (W < 2 || H < 2 || D < 2 || (long)W * H * D != s->n_nodes ||
      (long)(W - 1) * (H - 1) * (D - 1) != s->n_elems) {
    ufem_set_error("coarse: %d x %d x %d nodes do not match the mesh", W, H, D);
    return AF_ERR_SIZ;
  }
  if (cx < 1 || cy < 1 || cz < 1 || (W - 1) % cx || (H - 1) % cy || (D - 1) % cz ||
      cx * cy * cz > UGK_CC_MAX_CELL) {
    ufem_set_error("coarse: factors %d %d %d must divide the element counts (at most %d per cell)", cx, cy, cz,
                   UGK_CC_MAX_CELL);
    return AF_ERR_VAL;
  }
  if ((rv = check_box(s, W, H, D, h)) != AF_OK) return rv;
  C = (ugk_coarse *)calloc(1, sizeof(ugk_coarse));
  if (!C) return AF_ERR_MEM;
  C->W = W; C->H = H; C->D = D; C->cx = cx; C->cy = cy; C->cz = cz;
  C->Wc = (W - 1) / cx + 1; C->Hc = (H - 1) / cy + 1; C->Dc = (D - 1) / cz + 1;
  C->ncell = (long)(C->Wc - 1) * (C->Hc - 1) * (C->Dc - 1);
  nn = (long)C->Wc * C->Hc * C->Dc;
  C->nc = 3 * nn;
  C->ncp = (C->nc + 63) / 64 * 64;
  if (C->ncp > 16384) {   /* the dense inverse (n^3) would outweigh the solve */
    ufem_set_error("coarse: %ld coarse DOFs (at most 16384); use larger factors", C->nc);
    free(C);
    return AF_ERR_SIZ;
  }
  C->d_E = s->d_E;
  C->d_nu = s->d_nu;
  rv = AF_ERR_MEM;
  G = (double *)malloc(sizeof(double) * (size_t)cx * cy * cz * 2 * 576);
  fm = (double *)malloc(sizeof(double) * (size_t)n);
  cm = (double *)malloc(sizeof(double) * (size_t)C->nc);
  if (!G || !fm || !cm) goto fail;
  if ((rv = galerkin_tables(cx, cy, cz, h, G)) != AF_OK) goto fail;
  for (i = 0; i < n; i++) fm[i] = 1.0;
  for (i = 0; i < s->n_fixed; i++) fm[s->fixed_host[i]] = 0.0;
  for (i = 0; i < nn; i++) {   /* a coarse DOF follows its coincident fine DOF */
    const long a = i % C->Wc, b = (i / C->Wc) % C->Hc, c = i / ((long)C->Wc * C->Hc);
    const long f = a * cx + (long)W * (b * cy + (long)H * c * cz);
    int d;
    for (d = 0; d < 3; d++) cm[3 * i + d] = fm[3 * f + d];
  }
  rv = AF_ERR;
  C->d_G = (double *)ugk_malloc(sizeof(double) * (size_t)cx * cy * cz * 2 * 576);
  C->d_KC = (double *)ugk_malloc(sizeof(double) * (size_t)C->ncell * 576);
  C->d_A = (double *)ugk_malloc(sizeof(double) * (size_t)C->ncp * (size_t)C->ncp);
  C->d_piv = (double *)ugk_malloc(sizeof(double) * UGK_GJ_BLOCK * UGK_GJ_BLOCK);
  C->d_R = (double *)ugk_malloc(sizeof(double) * UGK_GJ_BLOCK * (size_t)C->ncp);
  C->d_C = (double *)ugk_malloc(sizeof(double) * UGK_GJ_BLOCK * (size_t)C->ncp);
  C->d_cpart = (double *)ugk_malloc(sizeof(double) * 24 * (size_t)C->ncell);
  C->d_g = (double *)ugk_malloc(sizeof(double) * (size_t)C->ncp);
  C->d_y = (double *)ugk_malloc(sizeof(double) * (size_t)C->ncp);
  C->d_fm = (double *)ugk_malloc(sizeof(double) * (size_t)n);
  C->d_cm = (double *)ugk_malloc(sizeof(double) * (size_t)C->nc);
  C->d_zc = (double *)ugk_malloc(sizeof(double) * (size_t)n);
  if (!C->d_G || !C->d_KC || !C->d_A || !C->d_piv || !C->d_R || !C->d_C || !C->d_cpart || !C->d_g || !C->d_y ||
      !C->d_fm || !C->d_cm || !C->d_zc) {
    gpu_fail("coarse: allocation");
    goto fail;
  }
  if (ugk_h2d(C->d_G, G, sizeof(double) * (size_t)cx * cy * cz * 2 * 576, s->stream) ||
      ugk_h2d(C->d_fm, fm, sizeof(double) * (size_t)n, s->stream) ||
      ugk_h2d(C->d_cm, cm, sizeof(double) * (size_t)C->nc, s->stream) || ugk_sync(s->stream)) {
    gpu_fail("coarse: upload");
    goto fail;
  }
  free(G); free(fm); free(cm);
  s->w.cc = C;
  return AF_OK;
fail:
  free(G); free(fm); free(cm);
  ufem_coarse_free(C);
  return rv;
}

int ufem_coarse_build(ugk_coarse *C, void *stream)
{
  if (ugk_cc_galerkin(C, stream) || ugk_cc_assemble(C, stream) ||
      ugk_dense_spd_inverse(C->d_A, C->ncp, C->d_piv, C->d_R, C->d_C, stream))
    return gpu_fail("coarse matrix");
  return AF_OK;
}

long ufem_sys_coarse_size(const ufem_sys *s) { return s->w.cc ? s->w.cc->nc : 0; }

int ufem_sys_coarse_matrix(ufem_sys *s, int inverse, double *A)
{
  ugk_coarse *C = s->w.cc;
  long i;
  if (!C) { ufem_set_error("coarse: ufem_sys_set_coarse first"); return AF_ERR; }
  if (ugk_cc_galerkin(C, s->stream) || ugk_cc_assemble(C, s->stream) ||
      (inverse && ugk_dense_spd_inverse(C->d_A, C->ncp, C->d_piv, C->d_R, C->d_C, s->stream)))
    return gpu_fail("coarse matrix");
  for (i = 0; i < C->nc; i++)
    if (ugk_d2h(A + i * C->nc, C->d_A + i * C->ncp, sizeof(double) * (size_t)C->nc, s->stream))
      return gpu_fail("coarse matrix read-back");
  return AF_OK;
}

int ufem_gpu_dense_spd_inverse(long n, const double *A, double *Ainv)
{
  const long np = (n + 63) / 64 * 64;
  double *dA = NULL, *piv = NULL, *R = NULL, *Cb = NULL, *h = NULL;
  void *strm = NULL;
  long i, j;
  int rv = AF_ERR;
  if (n <= 0) return AF_ERR_VAL;
  {
    char name[64];
    int ncu;
    long hbm;
    if (ufem_gpu_device_info(name, (int)sizeof(name), &ncu, &hbm) != AF_OK) return AF_ERR;
  }
  h = (double *)calloc((size_t)np * (size_t)np, sizeof(double));
  if (!h) return AF_ERR_MEM;
  for (i = 0; i < np; i++)
    for (j = 0; j < np; j++) h[i * np + j] = (i < n && j < n) ? A[i * n + j] : (i == j ? 1.0 : 0.0);
  strm = ugk_stream_create();
  dA = (double *)ugk_malloc(sizeof(double) * (size_t)np * (size_t)np);
  piv = (double *)ugk_malloc(sizeof(double) * UGK_GJ_BLOCK * UGK_GJ_BLOCK);
  R = (double *)ugk_malloc(sizeof(double) * UGK_GJ_BLOCK * (size_t)np);
  Cb = (double *)ugk_malloc(sizeof(double) * UGK_GJ_BLOCK * (size_t)np);
  if (!strm || !dA || !piv || !R || !Cb || ugk_h2d(dA, h, sizeof(double) * (size_t)np * (size_t)np, strm) ||
      ugk_dense_spd_inverse(dA, np, piv, R, Cb, strm) ||
      ugk_d2h(h, dA, sizeof(double) * (size_t)np * (size_t)np, strm)) {
    gpu_fail("dense inverse");
    goto out;
  }
  for (i = 0; i < n; i++) memcpy(Ainv + i * n, h + i * np, sizeof(double) * (size_t)n);
  rv = AF_OK;
out:
  ugk_free(dA); ugk_free(piv); ugk_free(R); ugk_free(Cb);
  ugk_stream_destroy(strm);
  free(h);
  return rv;
}
