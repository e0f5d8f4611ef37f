/*
 * ufem_ssor.c -- colour order for CG with the SSOR preconditioner.
 *
 * femEqsCGwSSOR (fem/fem_eqs.c:999-1188) sweeps symmetric Gauss-Seidel in
 * natural row order, one row after another.  On the GPU the block rows are
 * coloured greedily in natural order (a row takes the smallest colour none
 * of its already coloured neighbours has) and renumbered colour by colour,
 * natural order kept within a colour; a sweep is then one launch per colour
 * over a contiguous row range.  The colour-permuted matrix Kc = P A P^T is
 * a second SELL-64 copy whose pattern is built here once per matrix pattern
 * (A's pattern is read back from the device) and whose values are gathered
 * from A on the device at every solve.
 */
#define _POSIX_C_SOURCE 200809L
#include <stdlib.h>
#include <string.h>

#include "ufem_fem_math.h"
#include "ufem_internal.h"

#define SSOR_MAX_COLORS 256

void ufem_ssor_free(ufem_ssor *S)
{
  if (!S) return;
  ugk_free(S->Kc.d_slen); ugk_free(S->Kc.d_soff); ugk_free(S->Kc.d_rowlen); ugk_free(S->Kc.d_cols);
  ugk_free(S->Kc.d_vals); ugk_free(S->d_perm); ugk_free(S->d_dpos);
  free(S->cb);
  free(S);
}

static int cmp_int(const void *a, const void *b)
{
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

static int build(const ugk_mat *K, ufem_ssor *S, void *stream)
{
  const long nb = K->n_brows, nsl = K->n_slices, tot = K->total_slots;
  int *slen = NULL, *rowlen = NULL, *cols = NULL, *color = NULL, *perm = NULL, *inv = NULL, *ncols = NULL;
  int *nslen = NULL, *nrowlen = NULL, *row = NULL, *dpos = NULL, maxlen = 1;
  long *soff = NULL, *nsoff = NULL, *cnt = NULL, i, ntot = 0;
  int rv = AF_ERR_MEM, ncol = 0;
  unsigned char used[SSOR_MAX_COLORS];

  slen = (int *)malloc(sizeof(int) * (size_t)nsl);
  soff = (long *)malloc(sizeof(long) * (size_t)nsl);
  rowlen = (int *)malloc(sizeof(int) * (size_t)nb);
  cols = (int *)malloc(sizeof(int) * (size_t)tot * 64);
  color = (int *)malloc(sizeof(int) * (size_t)nb);
  perm = (int *)malloc(sizeof(int) * (size_t)nb);
  inv = (int *)malloc(sizeof(int) * (size_t)nb);
  cnt = (long *)calloc(SSOR_MAX_COLORS + 1, sizeof(long));
  if (!slen || !soff || !rowlen || !cols || !color || !perm || !inv || !cnt) goto out;
  if (ugk_d2h(slen, K->d_slen, sizeof(int) * (size_t)nsl, stream) ||
      ugk_d2h(soff, K->d_soff, sizeof(long) * (size_t)nsl, stream) ||
      ugk_d2h(rowlen, K->d_rowlen, sizeof(int) * (size_t)nb, stream) ||
      ugk_d2h(cols, K->d_cols, sizeof(int) * (size_t)tot * 64, stream)) {
    ufem_set_error("SSOR: pattern download: %s", ugk_error());
    rv = AF_ERR;
    goto out;
  }
  /* greedy colouring in natural order */
  for (i = 0; i < nb; i++) {
    const long sl = i >> 6, ln = i & 63;
    int k, c;
    memset(used, 0, sizeof(used));
    for (k = 0; k < rowlen[i]; k++) {
      const long J = cols[(soff[sl] + k) * 64 + ln];
      if (J < i) used[color[J]] = 1;
    }
    for (c = 0; c < SSOR_MAX_COLORS && used[c]; c++) {}
    if (c == SSOR_MAX_COLORS) {
      ufem_set_error("SSOR: more than %d colours", SSOR_MAX_COLORS);
      rv = AF_ERR_VAL;
      goto out;
    }
    color[i] = c;
    cnt[c + 1]++;
    if (c + 1 > ncol) ncol = c + 1;
  }
  for (i = 0; i < ncol; i++) cnt[i + 1] += cnt[i];
  S->cb = (long *)malloc(sizeof(long) * (size_t)(ncol + 1));
  if (!S->cb) goto out;
  memcpy(S->cb, cnt, sizeof(long) * (size_t)(ncol + 1));
  for (i = 0; i < nb; i++) {             /* stable: natural order within a colour */
    const long at = cnt[color[i]]++;
    perm[at] = (int)i;
    inv[i] = (int)at;
  }
  /* pattern of Kc: row I' = old row perm[I'], columns renumbered and sorted */
  nslen = (int *)calloc((size_t)nsl, sizeof(int));
  nsoff = (long *)malloc(sizeof(long) * (size_t)nsl);
  nrowlen = (int *)malloc(sizeof(int) * (size_t)nb);
  if (!nslen || !nsoff || !nrowlen) goto out;
  for (i = 0; i < nb; i++) {
    nrowlen[i] = rowlen[perm[i]];
    if (nrowlen[i] > nslen[i >> 6]) nslen[i >> 6] = nrowlen[i];
  }
  for (i = 0; i < nsl; i++) {
    nsoff[i] = ntot;
    ntot += nslen[i];
  }
  for (i = 0; i < nb; i++)
    if (nrowlen[i] > maxlen) maxlen = nrowlen[i];
  ncols = (int *)malloc(sizeof(int) * (size_t)ntot * 64);
  row = (int *)malloc(sizeof(int) * (size_t)maxlen);
  dpos = (int *)malloc(sizeof(int) * (size_t)nb);
  if (!ncols || !row || !dpos) goto out;
  for (i = 0; i < nsl; i++) {
    int l;
    for (l = 0; l < 64; l++) {
      const long I = i * 64 + l;
      int k;
      if (I < nb) {
        const long o = perm[I], so = soff[o >> 6], lo = o & 63;
        for (k = 0; k < nrowlen[I]; k++) row[k] = inv[cols[(so + k) * 64 + lo]];
        qsort(row, (size_t)nrowlen[I], sizeof(int), cmp_int);
        dpos[I] = -1;            /* the diagonal block splits the row: earlier colours, later ones */
        for (k = 0; k < nrowlen[I]; k++)
          if (row[k] == (int)I) dpos[I] = k;
        if (dpos[I] < 0) {
          ufem_set_error("SSOR: row %ld has no diagonal block", I);
          rv = AF_ERR_VAL;
          goto out;
        }
        for (k = 0; k < nslen[i]; k++) ncols[(nsoff[i] + k) * 64 + l] = k < nrowlen[I] ? row[k] : (int)I;
      } else {
        for (k = 0; k < nslen[i]; k++) ncols[(nsoff[i] + k) * 64 + l] = 0;
      }
    }
  }
  memset(&S->Kc, 0, sizeof(S->Kc));
  S->Kc.bs = K->bs;
  S->Kc.n_brows = nb;
  S->Kc.n_slices = nsl;
  S->Kc.total_slots = ntot;
  S->Kc.nnzb = K->nnzb;
  S->Kc.gen = 1;
  S->Kc.d_slen = (int *)ugk_malloc(sizeof(int) * (size_t)nsl);
  S->Kc.d_soff = (long *)ugk_malloc(sizeof(long) * (size_t)nsl);
  S->Kc.d_rowlen = (int *)ugk_malloc(sizeof(int) * (size_t)nb);
  S->Kc.d_cols = (int *)ugk_malloc(sizeof(int) * (size_t)ntot * 64);
  S->Kc.d_vals = (double *)ugk_malloc(sizeof(double) * (size_t)ntot * 64 * K->bs * K->bs);
  S->d_perm = (int *)ugk_malloc(sizeof(int) * (size_t)nb);
  S->d_dpos = (int *)ugk_malloc(sizeof(int) * (size_t)nb);
  if (!S->Kc.d_slen || !S->Kc.d_soff || !S->Kc.d_rowlen || !S->Kc.d_cols || !S->Kc.d_vals || !S->d_perm ||
      !S->d_dpos) {
    ufem_set_error("SSOR: allocation: %s", ugk_error());
    goto out;
  }
  if (ugk_h2d(S->Kc.d_slen, nslen, sizeof(int) * (size_t)nsl, stream) ||
      ugk_h2d(S->Kc.d_soff, nsoff, sizeof(long) * (size_t)nsl, stream) ||
      ugk_h2d(S->Kc.d_rowlen, nrowlen, sizeof(int) * (size_t)nb, stream) ||
      ugk_h2d(S->Kc.d_cols, ncols, sizeof(int) * (size_t)ntot * 64, stream) ||
      ugk_h2d(S->d_perm, perm, sizeof(int) * (size_t)nb, stream) ||
      ugk_h2d(S->d_dpos, dpos, sizeof(int) * (size_t)nb, stream) || ugk_sync(stream)) {
    ufem_set_error("SSOR: upload: %s", ugk_error());
    rv = AF_ERR;
    goto out;
  }
  S->ncol = ncol;
  rv = AF_OK;
out:
  free(slen); free(soff); free(rowlen); free(cols); free(color); free(perm); free(inv); free(cnt);
  free(nslen); free(nsoff); free(nrowlen); free(ncols); free(row); free(dpos);
  return rv;
}

/* colour order for K (rebuilt when K or its pattern changed), refilled with
 * K's current values */
int ufem_ssor_prepare(const ugk_mat *K, ufem_ssor **pS, void *stream)
{
  ufem_ssor *S = *pS;
  int rv;
  if (!S || S->src != K || S->gen != K->gen || S->src_slots != K->total_slots || S->Kc.n_brows != K->n_brows) {
    ufem_ssor_free(S);
    *pS = NULL;
    S = (ufem_ssor *)calloc(1, sizeof(ufem_ssor));
    if (!S) return AF_ERR_MEM;
    rv = build(K, S, stream);
    if (rv != AF_OK) {
      ufem_ssor_free(S);
      return rv;
    }
    S->src = K;
    S->gen = K->gen;
    S->src_slots = K->total_slots;
    *pS = S;
  }
  if (ugk_ssor_permute_vals(K, &S->Kc, S->d_perm, stream)) {
    ufem_set_error("SSOR: values: %s", ugk_error());
    return AF_ERR;
  }
  return AF_OK;
}
