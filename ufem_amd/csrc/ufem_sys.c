/*
 * ufem_sys.c -- host C side of the device-resident linear-static solve:
 * DOF map and sparsity pattern, element colouring, loads/supports
 * bookkeeping, and the calls into the HIP kernels (ugk.h).
 *
 * Mirrors femSolve's sequence (fem/fem_sol.c:1655-1817) for brick meshes;
 * each function cites the reference step it replaces.
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ufem_fem_math.h"
#include "ufem_internal.h"

static __thread char g_msg[512];

void ufem_set_error(const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
  va_end(ap);
}

const char *ufem_gpu_last_error(void) { return g_msg; }

static int gpu_fail(const char *what)
{
  ufem_set_error("%s: %s", what, ugk_error());
  return AF_ERR;
}

static int g_inited = 0;
static int ensure_gpu(void)
{
  if (!g_inited) {
    if (ugk_init() != 0) return gpu_fail("no usable GPU");
    g_inited = 1;
  }
  return AF_OK;
}

int ufem_gpu_device_info(char *name, int name_len, int *n_cu, long *hbm)
{
  if (ensure_gpu() != AF_OK) return AF_ERR;
  if (ugk_device_info(n_cu, hbm, name, name_len) != 0) return gpu_fail("device info");
  return AF_OK;
}

static int cmp_int(const void *a, const void *b)
{
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

/* node adjacency of the brick mesh -> SELL-64 pattern of 3x3 blocks.
 * The scalar pattern equals the reference's: K_e is added in full for every
 * element (fem/fem_mloc.c:182-216), so row 3I+r holds every column 3J+c of
 * every node J sharing an element with node I, explicit zeros included. */
static int build_pattern(ufem_sys *s, const ufem_mesh *m)
{
  long nn = m->n_nodes, ne = m->n_elems, i, k;
  long *incp = NULL, *inc = NULL, *fill = NULL;
  int *mark = NULL, *nb = NULL, *rowlen = NULL, *slen = NULL, *cols = NULL;
  long *soff = NULL, *nbp = NULL;
  long nslices = (nn + 63) / 64, total = 0, nnzb = 0;
  int rv = AF_ERR_MEM, maxb = 0;

  incp = (long *)calloc((size_t)nn + 1, sizeof(long));
  fill = (long *)calloc((size_t)nn, sizeof(long));
  mark = (int *)malloc(sizeof(int) * (size_t)nn);
  rowlen = (int *)calloc((size_t)nn, sizeof(int));
  nbp = (long *)calloc((size_t)nn + 1, sizeof(long));
  slen = (int *)calloc((size_t)nslices, sizeof(int));
  soff = (long *)calloc((size_t)nslices, sizeof(long));
  if (!incp || !fill || !mark || !rowlen || !nbp || !slen || !soff) goto out;
  for (i = 0; i < ne * 8; i++) incp[m->enodes[i] + 1]++;
  for (i = 0; i < nn; i++) {
    if (incp[i + 1] == 0) {
      ufem_set_error("node %ld belongs to no element", i);
      rv = AF_ERR_VAL;
      goto out;
    }
    incp[i + 1] += incp[i];
  }
  inc = (long *)malloc(sizeof(long) * (size_t)(ne * 8));
  if (!inc) goto out;
  for (i = 0; i < ne; i++)
    for (k = 0; k < 8; k++) {
      long nd = m->enodes[i * 8 + k];
      inc[incp[nd] + fill[nd]++] = i;
    }
  for (i = 0; i < nn; i++) mark[i] = -1;
  /* first pass: row lengths */
  for (i = 0; i < nn; i++) {
    int c = 0;
    for (k = incp[i]; k < incp[i + 1]; k++) {
      long e = inc[k], a;
      for (a = 0; a < 8; a++) {
        long J = m->enodes[e * 8 + a];
        if (mark[J] != (int)i) { mark[J] = (int)i; c++; }
      }
    }
    rowlen[i] = c;
    nbp[i + 1] = nbp[i] + c;
    if (c > maxb) maxb = c;
  }
  nnzb = nbp[nn];
  nb = (int *)malloc(sizeof(int) * (size_t)nnzb);
  if (!nb) goto out;
  for (i = 0; i < nn; i++) mark[i] = -1;
  for (i = 0; i < nn; i++) {
    long c = nbp[i];
    for (k = incp[i]; k < incp[i + 1]; k++) {
      long e = inc[k], a;
      for (a = 0; a < 8; a++) {
        long J = m->enodes[e * 8 + a];
        if (mark[J] != (int)i) { mark[J] = (int)i; nb[c++] = (int)J; }
      }
    }
    qsort(nb + nbp[i], (size_t)rowlen[i], sizeof(int), cmp_int);
  }
  for (i = 0; i < nslices; i++) {
    int l, mx = 0;
    for (l = 0; l < 64; l++) {
      long I = i * 64 + l;
      if (I < nn && rowlen[I] > mx) mx = rowlen[I];
    }
    slen[i] = mx;
    soff[i] = total;
    total += mx;
  }
  cols = (int *)malloc(sizeof(int) * (size_t)total * 64);
  if (!cols) goto out;
  for (i = 0; i < nslices; i++) {
    int l, sl;
    for (l = 0; l < 64; l++) {
      long I = i * 64 + l;
      for (sl = 0; sl < slen[i]; sl++) {
        int c;
        if (I < nn) c = sl < rowlen[I] ? nb[nbp[I] + sl] : (int)I;
        else c = 0;
        cols[(soff[i] + sl) * 64 + l] = c;
      }
    }
  }
  s->K.bs = 3;
  s->K.gen = 1;
  s->K.n_brows = nn;
  s->K.n_slices = nslices;
  s->K.total_slots = total;
  s->K.nnzb = nnzb;
  s->maxb = maxb;
  if (nslices > 4L * UGK_RED_MAX_BLOCKS) { ufem_set_error("mesh too large (%ld nodes)", nn); rv = AF_ERR_BIG; goto out; }
  s->K.d_slen = (int *)ugk_malloc(sizeof(int) * (size_t)nslices);
  s->K.d_soff = (long *)ugk_malloc(sizeof(long) * (size_t)nslices);
  s->K.d_rowlen = (int *)ugk_malloc(sizeof(int) * (size_t)nn);
  s->K.d_cols = (int *)ugk_malloc(sizeof(int) * (size_t)total * 64);
  s->K.d_vals = (double *)ugk_malloc(sizeof(double) * (size_t)total * 64 * 9);
  if (!s->K.d_slen || !s->K.d_soff || !s->K.d_rowlen || !s->K.d_cols || !s->K.d_vals) {
    rv = gpu_fail("pattern allocation");
    goto out;
  }
  if (ugk_h2d(s->K.d_slen, slen, sizeof(int) * (size_t)nslices, s->stream) ||
      ugk_h2d(s->K.d_soff, soff, sizeof(long) * (size_t)nslices, s->stream) ||
      ugk_h2d(s->K.d_rowlen, rowlen, sizeof(int) * (size_t)nn, s->stream) ||
      ugk_h2d(s->K.d_cols, cols, sizeof(int) * (size_t)total * 64, s->stream) ||
      ugk_sync(s->stream)) {
    rv = gpu_fail("pattern upload");
    goto out;
  }
  rv = AF_OK;
out:
  free(incp); free(inc); free(fill); free(mark); free(nb); free(rowlen); free(nbp);
  free(slen); free(soff); free(cols);
  return rv;
}

/* Greedy first-fit colouring in element order: elements sharing a node get
 * different colours, so the scatter of one colour never updates a block of K
 * twice.  On a structured brick grid this yields the 8 parity classes. */
static int color_elements(ufem_sys *s, const ufem_mesh *m)
{
  long nn = m->n_nodes, ne = m->n_elems, i, k;
  long *incp = NULL, *inc = NULL, *fill = NULL, *cnt = NULL;
  int *color = NULL, *list = NULL;
  unsigned char *used = NULL;
  int ncol = 0, rv = AF_ERR_MEM;
  const int MAXC = 256;
  incp = (long *)calloc((size_t)nn + 1, sizeof(long));
  fill = (long *)calloc((size_t)nn, sizeof(long));
  color = (int *)malloc(sizeof(int) * (size_t)ne);
  used = (unsigned char *)calloc(MAXC, 1);
  if (!incp || !fill || !color || !used) goto out;
  for (i = 0; i < ne * 8; i++) incp[m->enodes[i] + 1]++;
  for (i = 0; i < nn; i++) incp[i + 1] += incp[i];
  inc = (long *)malloc(sizeof(long) * (size_t)(ne * 8));
  if (!inc) goto out;
  for (i = 0; i < ne; i++)
    for (k = 0; k < 8; k++) {
      long nd = m->enodes[i * 8 + k];
      inc[incp[nd] + fill[nd]++] = i;
    }
  for (i = 0; i < ne; i++) {
    int c;
    memset(used, 0, MAXC);
    for (k = 0; k < 8; k++) {
      long nd = m->enodes[i * 8 + k], t;
      for (t = incp[nd]; t < incp[nd + 1]; t++) {
        long f = inc[t];
        if (f < i) used[color[f]] = 1;   /* only already-coloured elements */
      }
    }
    for (c = 0; c < MAXC && used[c]; c++) {}
    if (c == MAXC) { ufem_set_error("more than %d element colours", MAXC); rv = AF_ERR_BIG; goto out; }
    color[i] = c;
    if (c + 1 > ncol) ncol = c + 1;
  }
  cnt = (long *)calloc((size_t)ncol + 1, sizeof(long));
  list = (int *)malloc(sizeof(int) * (size_t)ne);
  s->color_off = (long *)calloc((size_t)ncol + 1, sizeof(long));
  if (!cnt || !list || !s->color_off) goto out;
  for (i = 0; i < ne; i++) s->color_off[color[i] + 1]++;
  for (i = 0; i < ncol; i++) s->color_off[i + 1] += s->color_off[i];
  for (i = 0; i < ne; i++) list[s->color_off[color[i]] + cnt[color[i]]++] = (int)i;
  s->n_colors = ncol;
  s->d_elist = (int *)ugk_malloc(sizeof(int) * (size_t)ne);
  if (!s->d_elist || ugk_h2d(s->d_elist, list, sizeof(int) * (size_t)ne, s->stream) || ugk_sync(s->stream)) {
    rv = gpu_fail("colour upload");
    goto out;
  }
  rv = AF_OK;
out:
  free(incp); free(inc); free(fill); free(cnt); free(color); free(list); free(used);
  return rv;
}

long ufem_sys_ndof(const ufem_sys *s) { return s->n; }
long ufem_sys_nnz(const ufem_sys *s) { return s->K.nnzb * 9; }
long ufem_sys_ncolors(const ufem_sys *s) { return s->n_colors; }
int ufem_mat_f32(ugk_mat *K, int rem, void *stream)
{
  if (!K->d_vals32) {
    K->d_vals32 = (float *)ugk_malloc(sizeof(float) * (size_t)K->total_slots * 64 * K->bs * K->bs);
    if (!K->d_vals32) return gpu_fail("fp32 matrix");
  }
  if (!rem && K->d_rem16) {
    ugk_free(K->d_rem16);
    K->d_rem16 = NULL;
  }
  if (rem && !K->d_rem16) {
    K->d_rem16 = (unsigned short *)ugk_malloc(sizeof(unsigned short) * (size_t)K->total_slots * 64 * K->bs * K->bs);
    if (!K->d_rem16) return gpu_fail("fp32 matrix remainders");
  }
  if (K->d_corr32 && getenv("UFEM_NO_CORR32")) {   /* plain FP32 copy, for comparisons */
    ugk_free(K->d_corr32);
    K->d_corr32 = NULL;
  }
  if (K->bs == 3 && !K->d_corr32 && !getenv("UFEM_NO_CORR32")) {
    K->d_corr32 = (float *)ugk_malloc(sizeof(float) * 9 * (size_t)K->n_brows);
    if (!K->d_corr32) return gpu_fail("fp32 matrix correction");
  }
  if (ugk_vals_to_f32(K, stream)) return gpu_fail("fp32 matrix");
  return AF_OK;
}

long ufem_sys_nreact(const ufem_sys *s) { return s->n_react; }

void ufem_sys_destroy(ufem_sys *s)
{
  int i;
  if (!s) return;
  if (s->stream) ugk_sync(s->stream);
  ugk_free(s->K.d_slen); ugk_free(s->K.d_soff); ugk_free(s->K.d_rowlen);
  ugk_free(s->K.d_cols); ugk_free(s->K.d_vals); ugk_free(s->K.d_vals32); ugk_free(s->K.d_corr32); ugk_free(s->K.d_rem16);
  ugk_free(s->d_x); ugk_free(s->d_y); ugk_free(s->d_z); ugk_free(s->d_E); ugk_free(s->d_nu);
  ugk_free(s->d_enodes); ugk_free(s->d_elist); ugk_free(s->d_fixed);
  ugk_free(s->d_saved_rows); ugk_free(s->d_saved_cols); ugk_free(s->d_Rfix);
  ugk_free(s->d_F); ugk_free(s->d_u);
  ufem_work_free(&s->w);
  for (i = 0; i < 6; i++) ugk_event_destroy(s->ev[i]);
  ugk_stream_destroy(s->stream);
  free(s->color_off); free(s->nl_node); free(s->nl_type); free(s->nl_dir); free(s->nl_val);
  free(s->F_host); free(s->F_load); free(s->fixed_host); free(s->react_node); free(s->react_dof);
  free(s->react_fix); free(s->react_init); free(s->fix_val);
  free(s->node_id); free(s->elem_id);
  free(s);
}

/* fem_dofs + femKhit + fem_sol_alloc + fem_sol_res_alloc
 * (fem/fem_sol.c:1263-1391, :485, :639-742; fem/fem_mloc.c:83-116) */
ufem_sys *ufem_sys_create(const ufem_mesh *m)
{
  ufem_sys *s;
  long i, ne, nn;
  double *tmp = NULL;
  int *itmp = NULL;
  if (ensure_gpu() != AF_OK) return NULL;
  if (!m || m->n_nodes < 1 || m->n_elems < 1) { ufem_set_error("empty mesh"); return NULL; }
  ne = m->n_elems;
  nn = m->n_nodes;
  for (i = 0; i < ne * 8; i++)
    if (m->enodes[i] < 0 || m->enodes[i] >= nn) { ufem_set_error("element node out of range"); return NULL; }
  for (i = 0; i < m->n_nl; i++) {
    if (m->nl_node[i] < 0 || m->nl_node[i] >= nn || m->nl_dir[i] < 1 || m->nl_dir[i] > 3) {
      ufem_set_error("load/support %ld out of range", i);
      return NULL;
    }
  }
  s = (ufem_sys *)calloc(1, sizeof(ufem_sys));
  if (!s) return NULL;
  s->n_nodes = nn;
  s->n_elems = ne;
  s->n = 3 * nn;
  s->stream = ugk_stream_create();
  for (i = 0; i < 6; i++) s->ev[i] = ugk_event_create();
  if (!s->stream || !s->ev[5]) { gpu_fail("stream"); goto fail; }
  if (build_pattern(s, m) != AF_OK) goto fail;
  if (color_elements(s, m) != AF_OK) goto fail;

  s->d_x = (double *)ugk_malloc(sizeof(double) * (size_t)nn);
  s->d_y = (double *)ugk_malloc(sizeof(double) * (size_t)nn);
  s->d_z = (double *)ugk_malloc(sizeof(double) * (size_t)nn);
  s->d_E = (double *)ugk_malloc(sizeof(double) * (size_t)ne);
  s->d_nu = (double *)ugk_malloc(sizeof(double) * (size_t)ne);
  s->d_enodes = (int *)ugk_malloc(sizeof(int) * (size_t)ne * 8);
  s->d_F = (double *)ugk_malloc(sizeof(double) * (size_t)s->n);
  s->d_u = (double *)ugk_malloc(sizeof(double) * (size_t)s->n);
  itmp = (int *)malloc(sizeof(int) * (size_t)ne * 8);
  if (!s->d_x || !s->d_y || !s->d_z || !s->d_E || !s->d_nu || !s->d_enodes || !s->d_F || !s->d_u || !itmp) {
    gpu_fail("mesh allocation");
    goto fail;
  }
  for (i = 0; i < ne * 8; i++) itmp[i] = (int)m->enodes[i];
  if (ugk_h2d(s->d_x, m->x, sizeof(double) * (size_t)nn, s->stream) ||
      ugk_h2d(s->d_y, m->y, sizeof(double) * (size_t)nn, s->stream) ||
      ugk_h2d(s->d_z, m->z, sizeof(double) * (size_t)nn, s->stream) ||
      ugk_h2d(s->d_E, m->E, sizeof(double) * (size_t)ne, s->stream) ||
      ugk_h2d(s->d_nu, m->nu, sizeof(double) * (size_t)ne, s->stream) ||
      ugk_h2d(s->d_enodes, itmp, sizeof(int) * (size_t)ne * 8, s->stream) || ugk_sync(s->stream)) {
    gpu_fail("mesh upload");
    goto fail;
  }
  free(itmp);
  itmp = NULL;
  if (ufem_work_alloc(&s->w, s->n, 1, 1) != AF_OK) goto fail;

  /* loads/supports in input order; reaction slots per support entry */
  s->n_nl = m->n_nl;
  s->nl_node = (long *)malloc(sizeof(long) * (size_t)(m->n_nl + 1));
  s->nl_type = (long *)malloc(sizeof(long) * (size_t)(m->n_nl + 1));
  s->nl_dir = (long *)malloc(sizeof(long) * (size_t)(m->n_nl + 1));
  s->nl_val = (double *)malloc(sizeof(double) * (size_t)(m->n_nl + 1));
  s->F_host = (double *)calloc((size_t)s->n, sizeof(double));
  s->F_load = (double *)calloc((size_t)s->n, sizeof(double));
  if (!s->nl_node || !s->nl_type || !s->nl_dir || !s->nl_val || !s->F_host || !s->F_load) goto fail;
  if (m->n_nl > 0) {
    memcpy(s->nl_node, m->nl_node, sizeof(long) * (size_t)m->n_nl);
    memcpy(s->nl_type, m->nl_type, sizeof(long) * (size_t)m->n_nl);
    memcpy(s->nl_dir, m->nl_dir, sizeof(long) * (size_t)m->n_nl);
    memcpy(s->nl_val, m->nl_val, sizeof(double) * (size_t)m->n_nl);
  }
  for (i = 0; i < m->n_nl; i++)
    if (m->nl_type[i] == 1) s->n_react++;
  s->react_node = (long *)calloc((size_t)s->n_react + 1, sizeof(long));
  s->react_dof = (long *)calloc((size_t)s->n_react + 1, sizeof(long));
  s->react_fix = (long *)calloc((size_t)s->n_react + 1, sizeof(long));
  s->react_init = (double *)calloc((size_t)s->n_react + 1, sizeof(double));
  s->fixed_host = (int *)calloc((size_t)s->n_react + 1, sizeof(int));
  s->fix_val = (double *)calloc((size_t)s->n_react + 1, sizeof(double));
  if (!s->react_node || !s->react_dof || !s->react_fix || !s->react_init || !s->fixed_host || !s->fix_val)
    goto fail;
  {
    long r = 0;
    char *seen = (char *)calloc((size_t)s->n, 1);
    long *slot_of = (long *)malloc(sizeof(long) * (size_t)s->n);
    if (!seen || !slot_of) { free(seen); free(slot_of); goto fail; }
    for (i = 0; i < m->n_nl; i++) {
      long d;
      if (m->nl_type[i] != 1) continue;
      d = 3 * m->nl_node[i] + m->nl_dir[i] - 1;
      s->react_node[r] = m->nl_node[i];
      s->react_dof[r] = m->nl_dir[i];
      if (!seen[d]) {
        seen[d] = 1;
        slot_of[d] = s->n_fixed;
        s->fix_val[s->n_fixed] = m->nl_val[i];
        if (m->nl_val[i] != 0.0) s->any_prescribed = 1;
        s->fixed_host[s->n_fixed++] = (int)d;
      } else if (m->nl_val[i] != s->fix_val[slot_of[d]]) {
        ufem_set_error("DOF prescribed twice with different values (entry %ld)", i);
        free(seen);
        free(slot_of);
        goto fail;
      }
      s->react_fix[r] = slot_of[d];
      r++;
    }
    free(seen);
    free(slot_of);
  }
  if (s->n_fixed > 0) {
    s->d_fixed = (int *)ugk_malloc(sizeof(int) * (size_t)s->n_fixed);
    s->d_saved_rows = (double *)ugk_malloc(sizeof(double) * (size_t)s->n_fixed * s->maxb * 3);
    s->d_saved_cols = (int *)ugk_malloc(sizeof(int) * (size_t)s->n_fixed * s->maxb);
    s->d_Rfix = (double *)ugk_malloc(sizeof(double) * (size_t)s->n_fixed);
    if (!s->d_fixed || !s->d_saved_rows || !s->d_saved_cols || !s->d_Rfix ||
        ugk_h2d(s->d_fixed, s->fixed_host, sizeof(int) * (size_t)s->n_fixed, s->stream) ||
        ugk_sync(s->stream)) {
      gpu_fail("support upload");
      goto fail;
    }
  }
  (void)tmp;
  return s;
fail:
  free(itmp);
  ufem_sys_destroy(s);
  return NULL;
}

/* fem_fill_K(AF_NO), fem/fem_sol.c:897-989: K = sum of element matrices,
 * here one kernel launch per colour in a fixed order */
int ufem_sys_set_E(ufem_sys *s, const double *E)
{
  if (ugk_h2d(s->d_E, E, sizeof(double) * (size_t)s->n_elems, s->stream) || ugk_sync(s->stream))
    return gpu_fail("E upload");
  s->filled = 0;
  s->bc_applied = 0;
  return AF_OK;
}

/* new values on an unchanged mesh (node coordinates, E, nu, load and support
 * values; the connectivity and the load/support list stay), then the state
 * monte_linear_solver starts from (fem/fem_mont.c:470-475: K, F and u set to
 * zero before fem_fill_K): F and the reaction slots cleared, K to be refilled */
int ufem_sys_update(ufem_sys *s, const ufem_mesh *m)
{
  long i, r = 0;
  if (!m || m->n_nodes != s->n_nodes || m->n_elems != s->n_elems || m->n_nl != s->n_nl) {
    ufem_set_error("update: mesh sizes differ from the system's");
    return AF_ERR_SIZ;
  }
  for (i = 0; i < m->n_nl; i++)
    if (m->nl_node[i] != s->nl_node[i] || m->nl_type[i] != s->nl_type[i] || m->nl_dir[i] != s->nl_dir[i]) {
      ufem_set_error("update: load/support %ld differs from the system's", i);
      return AF_ERR_VAL;
    }
  /* prescribed value of each fixed DOF from its first support entry; a later
     entry of the same DOF must agree, as at creation */
  {
    char *seen = (char *)calloc((size_t)s->n_fixed + 1, 1);
    if (!seen) return AF_ERR_MEM;
    for (i = 0; i < m->n_nl; i++) {
      long f;
      if (m->nl_type[i] != 1) continue;
      f = s->react_fix[r++];
      if (!seen[f]) {
        seen[f] = 1;
        s->fix_val[f] = m->nl_val[i];
      } else if (m->nl_val[i] != s->fix_val[f]) {
        ufem_set_error("DOF prescribed twice with different values (entry %ld)", i);
        free(seen);
        return AF_ERR_VAL;
      }
    }
    free(seen);
  }
  s->any_prescribed = 0;
  for (i = 0; i < s->n_fixed; i++)
    if (s->fix_val[i] != 0.0) s->any_prescribed = 1;
  if (m->n_nl > 0) memcpy(s->nl_val, m->nl_val, sizeof(double) * (size_t)m->n_nl);
  if (ugk_h2d(s->d_x, m->x, sizeof(double) * (size_t)s->n_nodes, s->stream) ||
      ugk_h2d(s->d_y, m->y, sizeof(double) * (size_t)s->n_nodes, s->stream) ||
      ugk_h2d(s->d_z, m->z, sizeof(double) * (size_t)s->n_nodes, s->stream) ||
      ugk_h2d(s->d_E, m->E, sizeof(double) * (size_t)s->n_elems, s->stream) ||
      ugk_h2d(s->d_nu, m->nu, sizeof(double) * (size_t)s->n_elems, s->stream) || ugk_sync(s->stream))
    return gpu_fail("mesh upload");
  memset(s->F_host, 0, sizeof(double) * (size_t)s->n);
  memset(s->F_load, 0, sizeof(double) * (size_t)s->n);
  memset(s->react_init, 0, sizeof(double) * (size_t)(s->n_react + 1));
  s->filled = 0;
  s->loaded = 0;
  s->bc_applied = 0;
  return AF_OK;
}

int ufem_sys_fill_K(ufem_sys *s, int npts)
{
  int c;
  ugk_state st;
  if (npts != 2 && npts != 3) { ufem_set_error("npts must be 2 or 3"); return AF_ERR_VAL; }
  if (ugk_zero(s->w.st, sizeof(ugk_state), s->stream)) return gpu_fail("state reset");
  if (ugk_event_record(s->ev[0], s->stream)) return gpu_fail("event");
  if (ugk_zero(s->K.d_vals, sizeof(double) * (size_t)s->K.total_slots * 64 * 9, s->stream))
    return gpu_fail("zero K");
  for (c = 0; c < s->n_colors; c++) {
    long off = s->color_off[c], cnt = s->color_off[c + 1] - off;
    if (ugk_elem_stiff_scatter(&s->K, cnt, s->d_elist + off, s->d_x, s->d_y, s->d_z, s->d_enodes,
                               s->d_E, s->d_nu, npts, s->w.st, s->stream))
      return gpu_fail("element stiffness");
  }
  if (ugk_event_record(s->ev[1], s->stream)) return gpu_fail("event");
  if (ugk_d2h(&st, s->w.st, sizeof(st), s->stream)) return gpu_fail("state read");
  s->t_fill_ms = ugk_event_ms(s->ev[0], s->ev[1]);
  if (st.err == 1) { ufem_set_error("element with non-positive Jacobian"); return AF_ERR_VAL; }
  if (st.err == 3) { ufem_set_error("element block missing from the pattern"); return AF_ERR; }
  s->filled = 1;
  s->bc_applied = 0;
  return AF_OK;
}

/* fem_add_loads(0): nodal forces (type 3) into F in input order,
 * fem/fem_sol.c:1119-1189, fem/fem_loads.c:343-397.  F starts from zero, as
 * the reference's F does at every femSolve (fem_sol_alloc), so a second call
 * gives the same F; the loads alone are kept for add_disps, which rebuilds
 * the supported F from them after every fill_K. */
int ufem_sys_add_loads(ufem_sys *s)
{
  long i;
  memset(s->F_load, 0, sizeof(double) * (size_t)s->n);
  for (i = 0; i < s->n_nl; i++)
    if (s->nl_type[i] == 3) s->F_load[3 * s->nl_node[i] + s->nl_dir[i] - 1] += s->nl_val[i];
  memcpy(s->F_host, s->F_load, sizeof(double) * (size_t)s->n);
  s->loaded = 1;
  s->bc_applied = 0;
  if (ugk_h2d(s->d_F, s->F_host, sizeof(double) * (size_t)s->n, s->stream) || ugk_sync(s->stream))
    return gpu_fail("F upload");
  return AF_OK;
}

/* fem_add_disps(AF_YES, 0) -> femApplyDisp with Val = 0 for every support in
 * input order (fem/fem_loads.c:116-208): the reaction slot takes -F at the
 * DOF, F and u there become 0, the row and the column are zeroed and the
 * diagonal set to 1.  The device zeroes the whole column (the reference's
 * serial build zeroes it only inside a window of rows, fem/fem_math.c:603-619;
 * the solution is the same because u is 0 at the DOF throughout CG).  The
 * original rows of the fixed DOFs are kept for the reactions. */
int ufem_sys_add_disps(ufem_sys *s)
{
  long r;
  if (!s->filled) { ufem_set_error("fill_K first"); return AF_ERR; }
  if (s->bc_applied) return AF_OK;
  /* F and the reaction slots from the loads alone: a refill of K (new
     moduli, fill_K) is followed by a fresh application, never by a second
     correction on top of the first */
  memcpy(s->F_host, s->F_load, sizeof(double) * (size_t)s->n);
  memset(s->react_init, 0, sizeof(double) * (size_t)(s->n_react + 1));
  for (r = 0; r < s->n_react; r++) {
    long d = 3 * s->react_node[r] + s->react_dof[r] - 1;
    long first = r, q;
    for (q = 0; q < r; q++)   /* fem2IntPos: the first slot with this node/DOF */
      if (s->react_node[q] == s->react_node[r] && s->react_dof[q] == s->react_dof[r]) { first = q; break; }
    s->react_init[first] -= s->F_host[d];
    s->F_host[d] = 0.0;
  }
  if (s->any_prescribed) {
    /* prescribed non-zero values: the free rows' loads lose K(:, p) v_p, as
     * femApplyDisp's row loop (fem/fem_loads.c:210-224) does one column at a
     * time -- here one SpMV with the vector of values, before the supports
     * zero K (the full rows and columns, so that no coupling is counted twice) */
    double *v = (double *)calloc((size_t)s->n, sizeof(double));
    double *y = (double *)malloc(sizeof(double) * (size_t)s->n);
    long i, f;
    int bad;
    if (!v || !y) { free(v); free(y); return AF_ERR_MEM; }
    for (f = 0; f < s->n_fixed; f++) v[s->fixed_host[f]] = s->fix_val[f];
    bad = ugk_h2d(s->w.z, v, sizeof(double) * (size_t)s->n, s->stream) ||
          ugk_spmv(&s->K, s->w.z, s->w.q, s->stream) ||
          ugk_d2h(y, s->w.q, sizeof(double) * (size_t)s->n, s->stream);
    if (!bad)
      for (i = 0; i < s->n; i++) s->F_host[i] -= y[i];
    free(v);
    free(y);
    if (bad) return gpu_fail("prescribed displacements");
  }
  {
    long f;   /* the support rows solve 1 u_p = v_p after the zeroing */
    for (f = 0; f < s->n_fixed; f++) s->F_host[s->fixed_host[f]] = s->fix_val[f];
  }
  if (ugk_event_record(s->ev[2], s->stream)) return gpu_fail("event");
  if (s->n_fixed > 0) {
    if (ugk_save_rows(&s->K, s->n_fixed, s->d_fixed, s->maxb, s->d_saved_rows, s->d_saved_cols, s->stream) ||
        ugk_apply_zero_bc(&s->K, s->n_fixed, s->d_fixed, s->stream))
      return gpu_fail("supports");
  }
  if (ugk_event_record(s->ev[3], s->stream)) return gpu_fail("event");
  if (ugk_h2d(s->d_F, s->F_host, sizeof(double) * (size_t)s->n, s->stream) || ugk_sync(s->stream))
    return gpu_fail("F upload");
  s->t_bc_ms = ugk_event_ms(s->ev[2], s->ev[3]);
  s->bc_applied = 1;
  return AF_OK;
}

int ufem_sys_solve(ufem_sys *s, const ufem_solve_opts *o, ufem_solve_info *info)
{
  ufem_solve_info tmp;
  int rv;
  if (!info) info = &tmp;
  memset(info, 0, sizeof(*info));
  if (!s->bc_applied) { ufem_set_error("supports not applied"); return AF_ERR; }
  if (!o->warm_start && ugk_zero(s->d_u, sizeof(double) * (size_t)s->n, s->stream)) return gpu_fail("u reset");
  rv = ufem_solve_device(&s->K, s->n, s->d_F, s->d_u, &s->w, s->stream, s->ev[4], s->ev[5], o, info);
  s->t_solve_ms = info->solve_ms;
  return rv;
}

int ufem_sys_solve_noalloc(ufem_sys *s, const ufem_solve_opts *o, ufem_solve_info *info)
{
  ufem_solve_info tmp;
  int rv;
  if (!info) info = &tmp;
  memset(info, 0, sizeof(*info));
  if (!s->bc_applied) { ufem_set_error("supports not applied"); return AF_ERR; }
  if (!o->warm_start && ugk_zero(s->d_u, sizeof(double) * (size_t)s->n, s->stream)) return gpu_fail("u reset");
  rv = ufem_solve_device_noalloc(&s->K, s->n, s->d_F, s->d_u, s->stream, o, info);
  s->t_solve_ms = info->solve_ms;
  return rv;
}

/* fem_fill_K(AF_YES) + fem_fill_K_React (fem/fem_sol.c:844-871): the
 * reaction at a support is (K u) there minus the load that stood there,
 * with K the matrix before supports */
int ufem_sys_reactions(ufem_sys *s, double *R, long *rnode, long *rdof)
{
  double *rf;
  long r;
  if (!s->bc_applied) { ufem_set_error("supports not applied"); return AF_ERR; }
  if (s->n_fixed == 0) return AF_OK;
  rf = (double *)malloc(sizeof(double) * (size_t)s->n_fixed);
  if (!rf) return AF_ERR_MEM;
  if (ugk_reactions(&s->K, s->n_fixed, s->d_fixed, s->maxb, s->d_saved_rows, s->d_saved_cols, s->d_u,
                    NULL, s->d_Rfix, s->stream) ||
      ugk_d2h(rf, s->d_Rfix, sizeof(double) * (size_t)s->n_fixed, s->stream)) {
    free(rf);
    return gpu_fail("reactions");
  }
  for (r = 0; r < s->n_react; r++) {
    /* a repeated support entry keeps only its own initial value, like the
       reference's per-entry slots (fem/fem_sol.c:844-871) */
    if (R) R[r] = s->react_init[r] + rf[s->react_fix[r]];
    if (rnode) rnode[r] = s->react_node[r];
    if (rdof) rdof[r] = s->react_dof[r];
  }
  free(rf);
  return AF_OK;
}

/* strains and stresses at the Gauss points (results pass of fem_fill_K(AF_YES),
 * fem/fem_e009.c:417-535) for the current u: res [n_elems][npts^3][12] */
int ufem_sys_element_results(ufem_sys *s, int npts, double *res)
{
  const long ng = (long)npts * npts * npts;
  const size_t bytes = sizeof(double) * (size_t)(s->n_elems * ng * 12);
  double *d = NULL;
  int rv = AF_OK;
  if (npts != 2 && npts != 3) { ufem_set_error("npts must be 2 or 3"); return AF_ERR_VAL; }
  d = (double *)ugk_malloc(bytes);
  if (!d) return gpu_fail("results buffer");
  if (ugk_elem_results(s->n_elems, npts, s->d_x, s->d_y, s->d_z, s->d_enodes, s->d_E, s->d_nu, s->d_u, d,
                       s->w.st, s->stream) ||
      ugk_d2h(res, d, bytes, s->stream))
    rv = gpu_fail("element results");
  ugk_free(d);
  return rv;
}

/* average device time (ms) of the results pass over `reps` launches, into a
 * device buffer (for measurement); negative on error */
double ufem_sys_time_results(ufem_sys *s, int npts, int reps)
{
  const long ng = (long)npts * npts * npts;
  double *d, ms = -1.0;
  int r;
  if ((npts != 2 && npts != 3) || reps < 1) return -1.0;
  d = (double *)ugk_malloc(sizeof(double) * (size_t)(s->n_elems * ng * 12));
  if (!d) return -1.0;
  if (!ugk_elem_results(s->n_elems, npts, s->d_x, s->d_y, s->d_z, s->d_enodes, s->d_E, s->d_nu, s->d_u, d,
                        s->w.st, s->stream) &&
      !ugk_event_record(s->ev[0], s->stream)) {
    for (r = 0; r < reps; r++)
      if (ugk_elem_results(s->n_elems, npts, s->d_x, s->d_y, s->d_z, s->d_enodes, s->d_E, s->d_nu, s->d_u, d,
                           s->w.st, s->stream))
        break;
    if (r == reps && !ugk_event_record(s->ev[1], s->stream) && !ugk_sync(s->stream))
      ms = ugk_event_ms(s->ev[0], s->ev[1]) / reps;
  }
  ugk_free(d);
  return ms;
}

/* femWriteRes (fem/fem_io.c:648-734): the reference's result file for the
 * current u -- solution header, node IDs, the DOF map (KNOWN_DOFS = 7 per
 * node, ux, uy, uz numbered), u, reactions, and the element results of the
 * 27-point results pass in the element's result layout: 20 repeating values
 * per point (sigma, epsilon, 6 plastic strains, 2 more; zero past epsilon
 * for a linear material) in a field of 14 x 27 values per element
 * (e009_rvals, fem/fem_e009.c:825-828), so that only the first 378 values,
 * points 1-18 and most of 19, are kept (fem/fem_accs.c:396-483).  Node and
 * element IDs are 1-based positions. */
#define WR_KNOWN_DOFS 7
#define WR_RES_RP 20
#define WR_ELEM_VALS (14 * 27)
/* IDs for the result file (the input's node and element IDs); NULL resets
 * to 1-based positions */
int ufem_sys_set_ids(ufem_sys *s, const long *node_id, const long *elem_id)
{
  free(s->node_id);
  free(s->elem_id);
  s->node_id = s->elem_id = NULL;
  if (node_id) {
    if (!(s->node_id = (long *)malloc(sizeof(long) * (size_t)s->n_nodes))) return AF_ERR_MEM;
    memcpy(s->node_id, node_id, sizeof(long) * (size_t)s->n_nodes);
  }
  if (elem_id) {
    if (!(s->elem_id = (long *)malloc(sizeof(long) * (size_t)s->n_elems))) return AF_ERR_MEM;
    memcpy(s->elem_id, elem_id, sizeof(long) * (size_t)s->n_elems);
  }
  return AF_OK;
}

int ufem_sys_set_fbc(ufem_sys *s, int fbc)
{
  s->fbc = fbc ? 1 : 0;
  return AF_OK;
}

/* a zero-valued support of the -fbc elimination (|value| <= FEM_ZERO, the
 * test of fem_dofs, fem/fem_sol.c:1286) */
#define FBC_ZERO 1e-17
static void fbc_mask(const ufem_sys *s, unsigned char *drop)
{
  long f;
  memset(drop, 0, (size_t)s->n);
  for (f = 0; f < s->n_fixed; f++)
    if (fabs(s->fix_val[f]) <= FBC_ZERO) drop[s->fixed_host[f]] = 1;
}

/* the reference's DOF map nDOFfld (fem_dofs, fem/fem_sol.c:1263-1391) with
 * WR_KNOWN_DOFS entries per node: the 1-based K position of ux, uy, uz,
 * node-major, the rest 0; with -fbc the zero-valued supports get 0 and the
 * positions close up.  Returns the number of active DOFs (nDOFAct). */
long ufem_sys_dof_map(const ufem_sys *s, long *map)
{
  unsigned char *drop = NULL;
  long i, act = 0;
  int d;
  if (s->fbc) {
    if (!(drop = (unsigned char *)malloc((size_t)s->n))) return AF_ERR_MEM;
    fbc_mask(s, drop);
  }
  for (i = 0; i < s->n_nodes; i++)
    for (d = 0; d < WR_KNOWN_DOFS; d++) {
      long v = 0;
      if (d < 3 && !(drop && drop[3 * i + d])) v = ++act;
      if (map) map[i * WR_KNOWN_DOFS + d] = v;
    }
  free(drop);
  return act;
}

int ufem_sys_write_results(ufem_sys *s, const char *fname, long sol_id, long dep_on, double sim_num)
{
  const long nn = s->n_nodes, ne = s->n_elems, n = s->n, nr = s->n_react;
  double *u = NULL, *R = NULL, *res = NULL;
  long *rn = NULL, *rd = NULL, *map = NULL, i, e, act;
  FILE *f = NULL;
  int rv = AF_ERR_MEM, d;
  map = (long *)malloc(sizeof(long) * (size_t)nn * WR_KNOWN_DOFS);
  if (!map) goto out;
  act = ufem_sys_dof_map(s, map);
  if (act < 0) goto out;
  u = (double *)malloc(sizeof(double) * (size_t)n);
  R = (double *)malloc(sizeof(double) * (size_t)(nr > 0 ? nr : 1));
  rn = (long *)malloc(sizeof(long) * (size_t)(nr > 0 ? nr : 1));
  rd = (long *)malloc(sizeof(long) * (size_t)(nr > 0 ? nr : 1));
  res = (double *)malloc(sizeof(double) * (size_t)(ne * 27 * 12));
  if (!u || !R || !rn || !rd || !res) goto out;
  if ((rv = ufem_sys_get_u(s, u)) != AF_OK) goto out;
  if ((rv = ufem_sys_reactions(s, R, rn, rd)) != AF_OK) goto out;
  if ((rv = ufem_sys_element_results(s, 3, res)) != AF_OK) goto out;
  if (!(f = fopen(fname, "w"))) {
    ufem_set_error("cannot write %s", fname);
    rv = AF_ERR_IO;
    goto out;
  }
  fprintf(f, "S\n %li %li %e \n", sol_id, dep_on, sim_num);
  fprintf(f, " %li \n %li \n %li\n", act, (long)WR_KNOWN_DOFS * nn, nn);
  for (i = 0; i < nn; i++) fprintf(f, " %li\n", s->node_id ? s->node_id[i] : i + 1);
  for (i = 0; i < nn * WR_KNOWN_DOFS; i++) fprintf(f, " %li\n", map[i]);
  /* u at the active DOFs in K order (the DOFs -fbc leaves out are 0) */
  fprintf(f, " %li\n", act);
  for (i = 0; i < nn; i++)
    for (d = 0; d < 3; d++)
      if (map[i * WR_KNOWN_DOFS + d] > 0) fprintf(f, " %e \n", u[3 * i + d]);
  (void)n;
  /* the reference's reaction list leaves out supports of non-zero value
     (fem_sol_res_alloc, fem/fem_sol.c:676-722) */
  {
    long nz = 0;
    for (i = 0; i < nr; i++) nz += s->fix_val[s->react_fix[i]] == 0.0;
    fprintf(f, " %li \n", nz);
  }
  /* under -fbc the eliminated supports keep the reaction slot's initial 0:
     femApplyDisp and fem_fill_K_React skip DOFs outside K (fem/fem_loads.c:132,
     fem/fem_sol.c:862) */
  for (i = 0; i < nr; i++)
    if (s->fix_val[s->react_fix[i]] == 0.0)
      fprintf(f, " %li %li %li %e\n", rn[i], s->node_id ? s->node_id[rn[i]] : rn[i] + 1, rd[i],
              s->fbc ? 0.0 : R[i]);
  fprintf(f, " %li %li\n", ne, (long)WR_ELEM_VALS * ne);
  for (e = 0; e < ne; e++) fprintf(f, " %li \n", s->elem_id ? s->elem_id[e] : e + 1);
  for (e = 0; e < ne; e++) fprintf(f, " %li \n", (long)WR_ELEM_VALS * e);
  for (e = 0; e < ne; e++)
    for (i = 0; i < WR_ELEM_VALS; i++) {
      const long gp = i / WR_RES_RP, t = i % WR_RES_RP;
      fprintf(f, " %e \n", t < 12 ? res[(e * 27 + gp) * 12 + t] : 0.0);
    }
  rv = fclose(f) == 0 ? AF_OK : AF_ERR_IO;
  f = NULL;
  if (rv != AF_OK) ufem_set_error("cannot close %s", fname);
out:
  if (f) fclose(f);
  free(u); free(R); free(rn); free(rd); free(res); free(map);
  return rv;
}

int ufem_sys_get_u(ufem_sys *s, double *u)
{
  if (ugk_d2h(u, s->d_u, sizeof(double) * (size_t)s->n, s->stream)) return gpu_fail("u read");
  return AF_OK;
}
int ufem_sys_set_u(ufem_sys *s, const double *u)
{
  if (ugk_h2d(s->d_u, u, sizeof(double) * (size_t)s->n, s->stream) || ugk_sync(s->stream))
    return gpu_fail("u write");
  return AF_OK;
}
int ufem_sys_get_F(ufem_sys *s, double *F)
{
  if (ugk_d2h(F, s->d_F, sizeof(double) * (size_t)s->n, s->stream)) return gpu_fail("F read");
  return AF_OK;
}

int ufem_sys_get_K(ufem_sys *s, long *indptr, int *cols, double *vals)
{
  long nn = s->n_nodes, tot = s->K.total_slots, nsl = s->K.n_slices, I;
  int *hc = (int *)malloc(sizeof(int) * (size_t)tot * 64);
  int *hl = (int *)malloc(sizeof(int) * (size_t)nn);
  long *ho = (long *)malloc(sizeof(long) * (size_t)nsl);
  double *hv = vals ? (double *)malloc(sizeof(double) * (size_t)tot * 64 * 9) : NULL;
  long k = 0;
  if (!hc || !hl || !ho || (vals && !hv)) { free(hc); free(hl); free(ho); free(hv); return AF_ERR_MEM; }
  if (ugk_d2h(hc, s->K.d_cols, sizeof(int) * (size_t)tot * 64, s->stream) ||
      ugk_d2h(hl, s->K.d_rowlen, sizeof(int) * (size_t)nn, s->stream) ||
      ugk_d2h(ho, s->K.d_soff, sizeof(long) * (size_t)nsl, s->stream) ||
      (vals && ugk_d2h(hv, s->K.d_vals, sizeof(double) * (size_t)tot * 64 * 9, s->stream))) {
    free(hc); free(hl); free(ho); free(hv);
    return gpu_fail("K read");
  }
  indptr[0] = 0;
  for (I = 0; I < nn; I++) {
    int r, sl, c;
    long so = ho[I >> 6], ln = I & 63;
    for (r = 0; r < 3; r++) {
      for (sl = 0; sl < hl[I]; sl++) {
        int J = hc[(so + sl) * 64 + ln];
        for (c = 0; c < 3; c++) {
          if (cols) cols[k] = 3 * J + c;
          if (vals) vals[k] = hv[ugk_vpos3(so + sl, 3 * r + c, (int)ln)];
          k++;
        }
      }
      indptr[3 * I + r + 1] = k;
    }
  }
  free(hc); free(hl); free(ho); free(hv);
  return AF_OK;
}

int ufem_sys_spmv(ufem_sys *s, const double *x, double *y)
{
  if (ugk_h2d(s->w.z, x, sizeof(double) * (size_t)s->n, s->stream) ||
      ugk_spmv(&s->K, s->w.z, s->w.q, s->stream) ||
      ugk_d2h(y, s->w.q, sizeof(double) * (size_t)s->n, s->stream))
    return gpu_fail("spmv");
  return AF_OK;
}

int ufem_sys_spmv_f32(ufem_sys *s, const double *x, double *y, int rem)
{
  int rv = ufem_mat_f32(&s->K, rem, s->stream);
  if (rv != AF_OK) return rv;
  if (ugk_h2d(s->w.z, x, sizeof(double) * (size_t)s->n, s->stream) ||
      ugk_spmv_f32x64(&s->K, s->w.z, s->w.q, s->stream) ||
      ugk_d2h(y, s->w.q, sizeof(double) * (size_t)s->n, s->stream))
    return gpu_fail("spmv (fp32 copy)");
  return AF_OK;
}

/* Algorithmic bytes of one launch, counted on the real (unpadded) blocks:
 *   mode 0  y = K x, FP64:      (72 + 4) nnzb + 8 n (x) + 8 n (y)
 *   mode 1  y = K x, FP32:      (36 + 4) nnzb + 4 n + 4 n
 *   mode 2  fused PCG kernel:   (72 + 4) nnzb + 8 n (z) + 8 n (p) + 8 n (p') + 8 n (q)
 *   mode 3  q = K p, (p, q):    (72 + 4) nnzb + 8 n (p) + 8 n (q)
 *   mode 4  w = K u, (w, u):    (72 + 4) nnzb + 8 n (u) + 8 n (w)
 * (n = 3 n_nodes; each vector counted once although its entries are
 * gathered by up to 27 rows -- those repeats are served by L2/MALL). */
static double kernel_bytes(const ugk_mat *K, int mode)
{
  double nnzb = (double)K->nnzb, n = (double)K->n_brows * K->bs, b2 = (double)K->bs * K->bs;
  if (mode == 1) return nnzb * (b2 * 4.0 + 4.0) + 8.0 * n;
  if (mode == 2) return nnzb * (b2 * 8.0 + 4.0) + 32.0 * n;
  if (mode == 5) return 16.0 * n;                      /* device copy of one vector */
  if (mode >= 6) return (double)K->total_slots * 64.0 * (b2 * 8.0 + 4.0);   /* stored values + columns */
  return nnzb * (b2 * 8.0 + 4.0) + 16.0 * n;   /* modes 0, 3, 4: x read, y written */
}

static int launch_mode(ufem_sys *s, int mode)
{
  if (mode == 1) return ugk_spmv_f32(&s->K, s->w.z32, s->w.q32, s->stream);
  if (mode == 2) return ugk_pcg_iter_a(&s->K, s->w.z, s->w.p[0], s->w.p[1], s->w.q, s->w.st, s->w.part, s->stream);
  if (mode == 3) return ugk_pcg_iter_q(&s->K, s->w.p[0], s->w.q, s->w.st, s->w.part, s->stream);
  if (mode == 4) return ugk_cg1_iter_w(&s->K, s->w.z, s->w.q, s->w.st, s->w.part, s->stream);
  if (mode == 5) return ugk_d2d(s->w.q, s->w.z, sizeof(double) * (size_t)s->n, s->stream);
  if (mode == 6 || mode == 7) return ugk_stream_read(&s->K, mode == 7, s->w.part, s->stream);
  return ugk_spmv(&s->K, s->w.z, s->w.q, s->stream);
}

int ufem_sys_time_spmv(ufem_sys *s, int reps, int mode, double *ms, double *bytes)
{
  int i;
  float t;
  if (reps < 1) reps = 1;
  if (mode < 0 || mode > 7) {
    ufem_set_error("ufem_sys_time_spmv: unknown kernel mode %d (0..7)", mode);
    return AF_ERR_VAL;
  }
  if (mode == 1 && !s->K.d_vals32) {
    int rv = ufem_mat_f32(&s->K, 0, s->stream);
    if (rv != AF_OK) return rv;
  }
  if (mode >= 2) {   /* a running state (after a solve) so the kernel does its work */
    ugk_state h;
    memset(&h, 0, sizeof(h));
    h.iter = 1;
    h.rule = 1;
    h.tol = 0.0;
    h.maxit = 1L << 40;
    h.bb = 1.0;
    h.rz[0] = h.rz[1] = 1.0;
    if (ugk_h2d(s->w.st, &h, sizeof(h), s->stream)) return gpu_fail("state");
  }
  if (launch_mode(s, mode)) return gpu_fail("kernel");   /* warm-up */
  if (ugk_event_record(s->ev[4], s->stream)) return gpu_fail("event");
  for (i = 0; i < reps; i++)
    if (launch_mode(s, mode)) return gpu_fail("kernel");
  if (ugk_event_record(s->ev[5], s->stream)) return gpu_fail("event");
  t = ugk_event_ms(s->ev[4], s->ev[5]);
  if (mode >= 2) {
    ugk_state h;
    if (ugk_d2h(&h, s->w.st, sizeof(h), s->stream)) return gpu_fail("state");
    if (h.done) { ufem_set_error("timed kernel exited early (state %d)", h.done); return AF_ERR; }
  }
  if (ms) *ms = (double)t / reps;
  if (bytes) *bytes = kernel_bytes(&s->K, mode);
  return AF_OK;
}

void ufem_sys_timings(const ufem_sys *s, double t[3])
{
  t[0] = s->t_fill_ms;
  t[1] = s->t_bc_ms;
  t[2] = s->t_solve_ms;
}

int ufem_gpu_element_stiffness(long n, const double *xyz, const double *E, const double *nu, int npts,
                               double *Ke)
{
  void *st = NULL, *strm = NULL;
  double *dxyz = NULL, *dE = NULL, *dnu = NULL, *dK = NULL;
  ugk_state *dst = NULL, hst;
  int rv = AF_ERR;
  if (ensure_gpu() != AF_OK) return AF_ERR;
  if (n <= 0) return AF_OK;
  strm = ugk_stream_create();
  dxyz = (double *)ugk_malloc(sizeof(double) * (size_t)n * 24);
  dE = (double *)ugk_malloc(sizeof(double) * (size_t)n);
  dnu = (double *)ugk_malloc(sizeof(double) * (size_t)n);
  dK = (double *)ugk_malloc(sizeof(double) * (size_t)n * 576);
  dst = (ugk_state *)ugk_malloc(sizeof(ugk_state));
  if (!strm || !dxyz || !dE || !dnu || !dK || !dst) { gpu_fail("alloc"); goto out; }
  if (ugk_h2d(dxyz, xyz, sizeof(double) * (size_t)n * 24, strm) ||
      ugk_h2d(dE, E, sizeof(double) * (size_t)n, strm) || ugk_h2d(dnu, nu, sizeof(double) * (size_t)n, strm) ||
      ugk_elem_stiff_dense(n, dxyz, NULL, dE, dnu, npts, dK, dst, strm) ||
      ugk_d2h(Ke, dK, sizeof(double) * (size_t)n * 576, strm) || ugk_d2h(&hst, dst, sizeof(hst), strm)) {
    gpu_fail("element stiffness");
    goto out;
  }
  rv = hst.err ? AF_ERR_VAL : AF_OK;
  if (hst.err) ufem_set_error("element with non-positive Jacobian");
out:
  (void)st;
  ugk_free(dxyz); ugk_free(dE); ugk_free(dnu); ugk_free(dK); ugk_free(dst);
  ugk_stream_destroy(strm);
  return rv;
}
