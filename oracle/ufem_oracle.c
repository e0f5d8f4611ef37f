/*
 * ufem_oracle.c -- clean-room CPU restatement of uFEM 0.5.3's linear-static
 * solve path (brick element 9 -> row-slot assembly -> supports -> Jacobi
 * CG/BiCGStab -> reactions).
 *
 * TEST INFRASTRUCTURE ONLY.  The parity tests, the smoke check and the
 * benchmark's CPU-baseline leg load this library as the checker.  Nothing in
 * the product (ufem_amd/) links, loads or calls it.
 *
 * Written from the reference's documented behaviour and read-as-text study of
 * its sources; every function cites the reference lines it restates.  In
 * serial mode the arithmetic is performed in the reference's operation order
 * so that, compiled with -ffp-contract=off like the reference's -ansi build,
 * results agree bit for bit with the compiled reference (pinned by
 * tests/test_oracle_golden.py against fixtures written by the compiled
 * reference).
 *
 * Known, documented deviations from the reference:
 *  - the results pass computes reactions but does not re-add K_e into K
 *    (fem/fem_sol.c:812 re-localises K_e; the matrix is not used afterwards);
 *  - support application always uses the serial window branch of
 *    femMatSetZeroCol (fem/fem_math.c:603-619); the threaded branch
 *    (fem/fem_math.c:574-584) reads frompos[-1] for empty slots.
 */
#define _POSIX_C_SOURCE 200809L
#include "ufem_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define KNOWN_DOFS 7          /* fem/fem_elem.h:43 */
#define FEM_ZERO 1e-17        /* fem/fem.h:59 */
#define MAX_THREADS 8         /* AF_MAX_THREADS, fem/fem.h:64 */
#define UO_OK 0
#define UO_ERR (-1)
#define UO_ERR_VAL (-3)
#define UO_ERR_ZER (-13)
#define UO_ERR_BIG (-5)
#define UO_ERR_MEM (-4)

typedef struct {              /* row-slot sparse matrix, fem/fem_math.h:47-57 */
  long rows, len;
  long *pos;                  /* 1-based column, 0 = empty slot */
  double *data;
  long *frompos, *defpos;
} rs_mat;

struct uo_sys {
  uo_model m;
  int threads;
  long ndof;
  long *dofmap;               /* nDOFfld: KNOWN_DOFS per node, 0 or 1..ndof */
  long *krows;                /* K_rows */
  rs_mat K;
  double *F, *u;
  long nreact;
  long *rnode, *rdof;
  double *rval;
  pthread_mutex_t *rowmtx;
  double t[5];
};

static double now_s(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ---------------------------------------------------------------------- */
/* Element 9: 8-node isoparametric brick                                   */
/* ---------------------------------------------------------------------- */

/* reference-cube corners, fem/fem_e009.c:36-69 */
static const double CUBE[8][3] = {
  {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
  {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

/* Gauss table with the reference's 15-digit literals, fem/fem_e008.c:470-489 */
static void gauss_table(int npts, double pt[3], double wt[3])
{
  if (npts == 2) {
    pt[0] = -0.577350269189626; wt[0] = 1.0;
    pt[1] = 0.577350269189626;  wt[1] = 1.0;
  } else {
    pt[0] = -0.774596669241483; wt[0] = 0.555555555555556;
    pt[1] = 0.000000000000000;  wt[1] = 0.888888888888889;
    pt[2] = 0.774596669241483;  wt[2] = 0.555555555555556;
  }
}

/* trilinear shape-function derivative d N_node / d xi_deriv in the
 * reference's evaluation order, fem/fem_e009.c:105-142 (e009_p_a) */
static double shape_d(int deriv, int node, double x, double y, double z)
{
  double xv = CUBE[node][0], yv = CUBE[node][1], zv = CUBE[node][2];
  switch (deriv) {
    case 1: return 0.125 * xv * (1.0 + yv * y) * (1.0 + zv * z);
    case 2: return 0.125 * yv * (1.0 + xv * x) * (1.0 + zv * z);
    default: return 0.125 * zv * (1.0 + xv * x) * (1.0 + yv * y);
  }
}

/* Crout LU with implicit row scaling and partial pivoting followed by
 * unit-column back substitution (the textbook ludcmp/lubksb pair), as used
 * for the 3x3 Jacobian: femLUdecomp/femLUback/femLUinverse,
 * fem/fem_math.c:2285-2483.  a is row-major n x n, overwritten by its inverse. */
static int lu_inverse3(double a[3][3])
{
  const int n = 3;
  double vv[3], b[3][3], col[3], big, dum, sum, temp;
  long index[3];
  int i, j, k, imax = 0, ii, ip;

  for (i = 0; i < n; i++) {
    big = 0.0;
    for (j = 0; j < n; j++)
      if ((temp = fabs(a[i][j])) > big) big = temp;
    if (big == 0.0) return UO_ERR_VAL;
    vv[i] = 1.0 / big;
  }
  for (j = 0; j < n; j++) {
    for (i = 0; i < j; i++) {
      sum = a[i][j];
      for (k = 0; k < i; k++) sum -= a[i][k] * a[k][j];
      a[i][j] = sum;
    }
    big = 0.0;
    for (i = j; i < n; i++) {
      sum = a[i][j];
      for (k = 0; k < j; k++) sum -= a[i][k] * a[k][j];
      a[i][j] = sum;
      if ((dum = vv[i] * fabs(sum)) >= big) { big = dum; imax = i; }
    }
    if (j != imax) {
      for (k = 0; k < n; k++) { dum = a[imax][k]; a[imax][k] = a[j][k]; a[j][k] = dum; }
      vv[imax] = vv[j];
    }
    index[j] = imax;
    if (a[j][j] == 0.0) a[j][j] = 1.0e-20;          /* FEM_TINY_VAL */
    if (j != n - 1) {
      dum = 1.0 / a[j][j];
      for (i = j + 1; i < n; i++) a[i][j] = dum * a[i][j];
    }
  }
  for (j = 0; j < n; j++) {
    for (i = 0; i < n; i++) col[i] = 0.0;
    col[j] = 1.0;
    ii = -1;
    for (i = 0; i < n; i++) {
      ip = (int)index[i];
      sum = col[ip];
      col[ip] = col[i];
      if (ii >= 0) {
        for (k = ii; k <= i - 1; k++) sum -= a[i][k] * col[k];
      } else if (sum != 0.0) {
        ii = i;
      }
      col[i] = sum;
    }
    for (i = n - 1; i >= 0; i--) {
      sum = col[i];
      for (k = i + 1; k < n; k++) sum -= a[i][k] * col[k];
      col[i] = sum / a[i][i];
    }
    for (i = 0; i < n; i++) b[i][j] = col[i];
  }
  for (i = 0; i < n; i++)
    for (j = 0; j < n; j++) a[i][j] = b[i][j];
  return UO_OK;
}

/* isotropic 3D elasticity, strain order ex,ey,ez,gyz,gzx,gxy:
 * femD_3D_iso, fem/fem_pl3d.c:32-58 */
static void iso_D(double E, double nu, double D[6][6])
{
  double d1 = (E * (1.0 - nu)) / ((1.0 + nu) * (1.0 - 2.0 * nu));
  double d2 = (E * nu) / ((1.0 + nu) * (1.0 - 2.0 * nu));
  double d3 = E / (2.0 * (1 + nu));
  memset(D, 0, sizeof(double) * 36);
  D[0][0] = D[1][1] = D[2][2] = d1;
  D[0][1] = D[0][2] = D[1][0] = D[2][0] = D[1][2] = D[2][1] = d2;
  D[3][3] = D[4][4] = D[5][5] = d3;
}

/* per-Gauss-point kinematics: J (fem/fem_e009.c:189-205), det J
 * (fem/fem_e009.c:76-98), J^-1, and the 6x24 strain matrix H
 * (fem/fem_e009.c:207-259).  Returns det J * w, or a negative code. */
static int gp_kinematics(const double xyz[24], double x, double y, double z,
                         double wx, double wy, double wz, double H[6][24], double *mult)
{
  double J[3][3], detj;
  int i, j, n;
  for (i = 0; i < 3; i++) {
    for (j = 0; j < 3; j++) {
      double val = 0.0;
      for (n = 0; n < 8; n++) {
        double vi = shape_d(i + 1, n, x, y, z);
        vi *= xyz[3 * n + j];
        val += vi;
      }
      J[i][j] = 0.0 + val;   /* femMatSetZero then femMatAdd */
    }
  }
  detj = J[0][0] * J[1][1] * J[2][2] +
         J[1][0] * J[2][1] * J[0][2] +
         J[0][1] * J[1][2] * J[2][0] +
         -J[2][0] * J[1][1] * J[0][2]
         - J[0][0] * J[1][2] * J[2][1]
         - J[2][2] * J[1][0] * J[0][1];
  if (detj <= 0.0) return UO_ERR_VAL;
  *mult = detj * wx * wy * wz;
  if (lu_inverse3(J) != UO_OK) return UO_ERR_VAL;
  memset(H, 0, sizeof(double) * 6 * 24);
  for (n = 0; n < 8; n++) {
    double h11 = 0.0, h22 = 0.0, h33 = 0.0;
    for (j = 0; j < 3; j++) {
      h11 += J[0][j] * shape_d(j + 1, n, x, y, z);
      h22 += J[1][j] * shape_d(j + 1, n, x, y, z);
      h33 += J[2][j] * shape_d(j + 1, n, x, y, z);
    }
    H[0][3 * n + 0] = h11;
    H[1][3 * n + 1] = h22;
    H[2][3 * n + 2] = h33;
    H[5][3 * n + 0] = h22;  /* gxy */
    H[5][3 * n + 1] = h11;
    H[3][3 * n + 1] = h33;  /* gyz */
    H[3][3 * n + 2] = h22;
    H[4][3 * n + 0] = h33;  /* gzx */
    H[4][3 * n + 2] = h11;
  }
  return UO_OK;
}

/* K_e = sum_gp detJ w H^T D H, e009_stiff Mode=AF_NO, fem/fem_e009.c:261-635
 * (GP loops :369-373, HTD/K_i products :587-602 via femMatMatMult
 * fem/fem_math.c:1665-1700). */
int uo_element_stiffness(const double xyz[24], double E, double nu, int npts, double Ke[576])
{
  double pt[3], wt[3], H[6][24], HT[24][6], D[6][6], HTD[24][6], mult, val;
  int i, j, k, ii, jj, kk, rv;
  if (npts != 2 && npts != 3) return UO_ERR_VAL;
  gauss_table(npts, pt, wt);
  iso_D(E, nu, D);
  for (i = 0; i < 576; i++) Ke[i] = 0.0;
  for (i = 0; i < npts; i++) {
    for (j = 0; j < npts; j++) {
      for (k = 0; k < npts; k++) {
        rv = gp_kinematics(xyz, pt[i], pt[j], pt[k], wt[i], wt[j], wt[k], H, &mult);
        if (rv != UO_OK) return rv;
        for (ii = 0; ii < 24; ii++)
          for (jj = 0; jj < 6; jj++) HT[ii][jj] = H[jj][ii];
        for (ii = 0; ii < 24; ii++) {
          for (jj = 0; jj < 6; jj++) {
            val = 0;
            for (kk = 0; kk < 6; kk++) val += HT[ii][kk] * D[kk][jj];
            HTD[ii][jj] = val;
          }
        }
        for (ii = 0; ii < 24; ii++) {
          for (jj = 0; jj < 24; jj++) {
            val = 0;
            for (kk = 0; kk < 6; kk++) val += HTD[ii][kk] * H[kk][jj];
            Ke[ii * 24 + jj] += mult * val;
          }
        }
      }
    }
  }
  return UO_OK;
}

/* F_e = sum_gp detJ w H^T (D H u_e): e009_stiff Mode=AF_YES with the default
 * femTangentMatrix = AF_YES (fem/fem_sol.c:35) and a material-1 element, the
 * "else mtype" branch fem/fem_e009.c:503-535 and :571-577.  gp (optional):
 * the results that branch adds at each point (femAddEResVal, :521-535),
 * [npts^3][12] = sigma (RES_SX .. RES_SXY) then epsilon (RES_EX .. RES_EXY). */
static int element_forces(const double xyz[24], double E, double nu, int npts,
                          const double ue[24], double Fe[24], double *gp)
{
  double pt[3], wt[3], H[6][24], D[6][6], eps[6], dsig[6], sig[6], F0[24], mult, val;
  int i, j, k, ii, jj, rv;
  gauss_table(npts, pt, wt);
  iso_D(E, nu, D);
  for (ii = 0; ii < 24; ii++) Fe[ii] = 0.0;
  for (i = 0; i < npts; i++) {
    for (j = 0; j < npts; j++) {
      for (k = 0; k < npts; k++) {
        rv = gp_kinematics(xyz, pt[i], pt[j], pt[k], wt[i], wt[j], wt[k], H, &mult);
        if (rv != UO_OK) return rv;
        for (ii = 0; ii < 6; ii++) {          /* epsilon = H u_e */
          val = 0.0;
          for (jj = 0; jj < 24; jj++) val += ue[jj] * H[ii][jj];
          eps[ii] = val;
        }
        for (ii = 0; ii < 6; ii++) {          /* d_sigma = D_r epsilon */
          val = 0.0;
          for (jj = 0; jj < 6; jj++) val += eps[jj] * D[ii][jj];
          dsig[ii] = val;
        }
        for (ii = 0; ii < 6; ii++) {          /* sigma = D epsilon */
          val = 0.0;
          for (jj = 0; jj < 6; jj++) val += eps[jj] * D[ii][jj];
          sig[ii] = val;
        }
        (void)dsig;
        if (gp) {
          double *o = gp + 12L * ((i * npts + j) * npts + k);
          for (ii = 0; ii < 6; ii++) { o[ii] = sig[ii]; o[6 + ii] = eps[ii]; }
        }
        for (ii = 0; ii < 24; ii++) {         /* F_0 = H^T sigma */
          val = 0.0;
          for (jj = 0; jj < 6; jj++) val += sig[jj] * H[jj][ii];
          F0[ii] = val;
        }
        for (ii = 0; ii < 24; ii++) Fe[ii] += mult * F0[ii];
      }
    }
  }
  return UO_OK;
}

/* ---------------------------------------------------------------------- */
/* Row-slot sparse matrix, fem/fem_math.c:57-235                           */
/* ---------------------------------------------------------------------- */

static int rs_alloc(rs_mat *a, long rows, const long *rowdesc)
{
  long i, sum = 0;
  a->rows = rows;
  a->defpos = (long *)calloc((size_t)rows, sizeof(long));
  a->frompos = (long *)calloc((size_t)rows, sizeof(long));
  if (!a->defpos || !a->frompos) return UO_ERR_MEM;
  for (i = 0; i < rows; i++) {
    sum += rowdesc[i];
    a->defpos[i] = rowdesc[i];
    a->frompos[i] = sum - rowdesc[i];
  }
  a->len = sum;
  a->data = (double *)calloc((size_t)sum + 1, sizeof(double));
  a->pos = (long *)calloc((size_t)sum + 1, sizeof(long));
  if (!a->data || !a->pos) return UO_ERR_MEM;
  return UO_OK;
}

static void rs_free(rs_mat *a)
{
  free(a->defpos); free(a->frompos); free(a->data); free(a->pos);
  memset(a, 0, sizeof(*a));
}

/* femMatPutAdd, sparse branch, fem/fem_math.c:203-228 */
static int rs_putadd(rs_mat *a, long row, long col, double val, int add)
{
  long i, from = a->frompos[row - 1], to = from + a->defpos[row - 1];
  for (i = from; i < to; i++) {
    if (a->pos[i] == col) {
      if (add) a->data[i] += val; else a->data[i] = val;
      return UO_OK;
    }
    if (a->pos[i] == 0) {
      a->pos[i] = col;
      if (add) a->data[i] += val; else a->data[i] = val;
      return UO_OK;
    }
  }
  return UO_ERR_BIG;
}

/* femMatGet, sparse branch, fem/fem_math.c:156-166 */
static double rs_get(const rs_mat *a, long row, long col)
{
  long i, from = a->frompos[row - 1], to = from + a->defpos[row - 1];
  for (i = from; i < to; i++) {
    if (a->pos[i] == 0) break;
    if (a->pos[i] == col) return a->data[i];
  }
  return 0.0;
}

/* femMatSetZeroRow, fem/fem_math.c:535-555 */
static void rs_zero_row(rs_mat *a, long row)
{
  long i, from = a->frompos[row - 1], to = from + a->defpos[row - 1];
  for (i = from; i < to; i++) {
    if (a->pos[i] == 0) break;
    a->data[i] = 0;
  }
}

/* femMatSetZeroCol, the branch compiled into the default threaded build when
 * threads are off: the window of rows starts at the first column stored in
 * row Col and spans defpos[Col-1] rows, fem/fem_math.c:603-619 */
static void rs_zero_col(rs_mat *a, long Col)
{
  long i, j, ifrom, ito;
  ifrom = a->pos[a->frompos[Col - 1]] - 1;
  ito = a->pos[a->frompos[Col - 1]] + a->defpos[Col - 1] - 1;
  for (i = ifrom; i < ito; i++) {
    if (i < 0 || i >= a->rows) continue;   /* the reference reads past the
                                              last row here; those rows hold
                                              no such column */
    for (j = a->frompos[i]; j < a->frompos[i] + a->defpos[i]; j++)
      if (a->pos[j] == Col) a->data[j] = 0.0;
  }
}

/* ---------------------------------------------------------------------- */
/* fork-join kernels: contiguous row blocks, ldiv split with the remainder */
/* on the last thread (fem/fem_math.c:1045-1080 and siblings)              */
/* ---------------------------------------------------------------------- */

typedef struct {
  long from, to;
  const rs_mat *A;
  const double *x, *y;
  double *out;
  double acc;
  int op;
} th_job;

enum { OP_SPMV = 0, OP_DOT = 1, OP_NORM2 = 2, OP_MATNORM = 3 };

static void run_job(th_job *p)
{
  long i, j;
  double acc = 0.0;
  if (p->to <= 0) { p->acc = 0.0; return; }
  switch (p->op) {
    case OP_SPMV:   /* thFemMatVecMult, fem/fem_math.c:1321-1342 */
      for (i = p->from; i < p->to; i++) {
        double val = 0.0;
        for (j = p->A->frompos[i]; j < p->A->frompos[i] + p->A->defpos[i]; j++) {
          if (p->A->pos[j] <= 0) break;
          val += (p->A->data[j] * p->x[p->A->pos[j] - 1]);
        }
        p->out[i] = val;
      }
      break;
    case OP_DOT:    /* thFemVecVecMult, fem/fem_math.c:988-1009 */
      for (i = p->from; i < p->to; i++) acc += p->x[i] * p->y[i];
      break;
    case OP_NORM2:  /* thFemVecNorm (pow(x,2) == x*x exactly), fem/fem_math.c:1994-2015 */
      for (i = p->from; i < p->to; i++) acc += p->x[i] * p->x[i];
      break;
    case OP_MATNORM: /* thFemMatNorm, fem/fem_math.c:1819-1849 */
      for (i = p->from; i < p->to; i++) {
        double nr = 0.0;
        for (j = p->A->frompos[i]; j < p->A->frompos[i] + p->A->defpos[i]; j++) {
          if (p->A->pos[j] <= 0) break;
          nr += p->A->data[j] * p->A->data[j];
        }
        nr = sqrt(nr);
        if (nr > acc) acc = nr;
      }
      break;
  }
  p->acc = acc;
}

static void *th_entry(void *arg) { run_job((th_job *)arg); return NULL; }

static double fork_join(int threads, int op, long n, const rs_mat *A,
                        const double *x, const double *y, double *out)
{
  th_job jobs[MAX_THREADS];
  pthread_t tid[MAX_THREADS];
  long quot, rem, i;
  double acc = 0.0;
  if (threads <= 1) {
    th_job j0;
    j0.from = 0; j0.to = n; j0.A = A; j0.x = x; j0.y = y; j0.out = out; j0.op = op;
    run_job(&j0);
    return j0.acc;
  }
  quot = n / threads; rem = n % threads;
  if (quot < 1) { quot = 0; rem = n; }
  for (i = 0; i < threads; i++) {
    jobs[i].from = i * quot; jobs[i].to = (i + 1) * quot;
    if (i == threads - 1) jobs[i].to += rem;
    jobs[i].A = A; jobs[i].x = x; jobs[i].y = y; jobs[i].out = out; jobs[i].op = op;
    pthread_create(&tid[i], NULL, th_entry, &jobs[i]);
  }
  for (i = 0; i < threads; i++) pthread_join(tid[i], NULL);
  for (i = 0; i < threads; i++) {
    if (op == OP_MATNORM) { if (jobs[i].acc > acc) acc = jobs[i].acc; }
    else acc += jobs[i].acc;
  }
  return acc;
}

static void spmv(const uo_sys *s, const double *x, double *y)
{ fork_join(s->threads, OP_SPMV, s->K.rows, &s->K, x, NULL, y); }
static double dot(const uo_sys *s, const double *a, const double *b, long n)
{ return fork_join(s->threads, OP_DOT, n, NULL, a, b, NULL); }
static double vnorm(const uo_sys *s, const double *a, long n)
{ return sqrt(fork_join(s->threads, OP_NORM2, n, NULL, a, NULL, NULL)); }

/* ---------------------------------------------------------------------- */
/* model set-up: DOF numbering, row capacities, reaction slots             */
/* ---------------------------------------------------------------------- */

/* fem_dofs (fem/fem_sol.c:1263-1391) + femKhit (fem/fem_mloc.c:83-116) +
 * fem_sol_alloc (fem/fem_sol.c:485) + fem_sol_res_alloc (fem/fem_sol.c:639-742) */
uo_sys *uo_create(const uo_model *m, int threads)
{
  uo_sys *s;
  long i, j, k, act = 0, nl;
  double t0 = now_s();
  s = (uo_sys *)calloc(1, sizeof(uo_sys));
  if (!s) return NULL;
  s->m = *m;
  s->threads = threads < 1 ? 1 : (threads > MAX_THREADS ? MAX_THREADS : threads);
  s->dofmap = (long *)calloc((size_t)m->n_nodes * KNOWN_DOFS, sizeof(long));
  if (!s->dofmap) { uo_free(s); return NULL; }
  for (i = 0; i < m->n_elems; i++)
    for (j = 0; j < 8; j++) {
      long npos = m->enodes[8 * i + j];
      for (k = 0; k < 3; k++) {
        long d = npos * KNOWN_DOFS + k;
        if (s->dofmap[d] == 0) { s->dofmap[d] = 1; act++; }
      }
    }
  s->ndof = act;
  act = 1;
  for (i = 0; i < m->n_nodes * KNOWN_DOFS; i++)
    if (s->dofmap[i] != 0) s->dofmap[i] = act++;
  s->krows = (long *)calloc((size_t)s->ndof, sizeof(long));
  if (!s->krows) { uo_free(s); return NULL; }
  for (i = 0; i < m->n_elems; i++)
    for (j = 0; j < 8; j++)
      for (k = 0; k < 3; k++) {
        long p = s->dofmap[m->enodes[8 * i + j] * KNOWN_DOFS + k];
        if (p >= 1) s->krows[p - 1] += 24;
      }
  if (rs_alloc(&s->K, s->ndof, s->krows) != UO_OK) { uo_free(s); return NULL; }
  s->F = (double *)calloc((size_t)s->ndof, sizeof(double));
  s->u = (double *)calloc((size_t)s->ndof, sizeof(double));
  s->rowmtx = (pthread_mutex_t *)malloc(sizeof(pthread_mutex_t) * (size_t)s->ndof);
  if (!s->F || !s->u || !s->rowmtx) { uo_free(s); return NULL; }
  for (i = 0; i < s->ndof; i++) pthread_mutex_init(&s->rowmtx[i], NULL);
  nl = 0;
  for (i = 0; i < m->n_nl; i++)
    if (m->nl_type[i] == 1 && m->nl_val[i] == 0.0) nl++;
  s->nreact = nl;
  s->rnode = (long *)calloc((size_t)nl + 1, sizeof(long));
  s->rdof = (long *)calloc((size_t)nl + 1, sizeof(long));
  s->rval = (double *)calloc((size_t)nl + 1, sizeof(double));
  nl = 0;
  for (i = 0; i < m->n_nl; i++)
    if (m->nl_type[i] == 1 && m->nl_val[i] == 0.0) {
      s->rnode[nl] = m->nl_node[i]; s->rdof[nl] = m->nl_dir[i]; nl++;
    }
  s->t[0] = now_s() - t0;
  return s;
}

void uo_free(uo_sys *s)
{
  long i;
  if (!s) return;
  if (s->rowmtx) { for (i = 0; i < s->ndof; i++) pthread_mutex_destroy(&s->rowmtx[i]); }
  free(s->rowmtx); free(s->dofmap); free(s->krows); free(s->F); free(s->u);
  free(s->rnode); free(s->rdof); free(s->rval);
  rs_free(&s->K);
  free(s);
}

long uo_ndof(const uo_sys *s) { return s->ndof; }
long uo_nreact(const uo_sys *s) { return s->nreact; }

static void elem_xyz(const uo_model *m, long e, double xyz[24])
{
  int n;
  for (n = 0; n < 8; n++) {
    long p = m->enodes[8 * e + n];
    xyz[3 * n + 0] = m->x[p]; xyz[3 * n + 1] = m->y[p]; xyz[3 * n + 2] = m->z[p];
  }
}

static void elem_locvec(const uo_sys *s, long e, long lf[24])
{   /* femFillLocVec, fem/fem_mloc.c:121-154 */
  int n, k;
  for (n = 0; n < 8; n++)
    for (k = 0; k < 3; k++) {
      long p = s->dofmap[s->m.enodes[8 * e + n] * KNOWN_DOFS + k];
      lf[3 * n + k] = p < 1 ? 0 : p;
    }
}

typedef struct { uo_sys *s; long from, to; int npts; int rv; } fill_job;

/* fem_fill_part_of_K (fem/fem_sol.c:748-840) with femLocKM_e
 * (fem/fem_mloc.c:163-221): K(lFld[i], lFld[j]) += K_e(i,j), row i locked */
static void fill_part(fill_job *j)
{
  double xyz[24], Ke[576];
  long lf[24], e;
  int i, k, rv;
  uo_sys *s = j->s;
  j->rv = UO_OK;
  for (e = j->from; e < j->to; e++) {
    elem_xyz(&s->m, e, xyz);
    rv = uo_element_stiffness(xyz, s->m.E[e], s->m.nu[e], j->npts, Ke);
    if (rv != UO_OK) { j->rv = rv; return; }
    elem_locvec(s, e, lf);
    for (i = 0; i < 24; i++) {
      if (lf[i] == 0) continue;
      if (s->threads > 1) pthread_mutex_lock(&s->rowmtx[lf[i] - 1]);
      for (k = 0; k < 24; k++) {
        if (lf[i] * lf[k] == 0) continue;
        rv = rs_putadd(&s->K, lf[i], lf[k], Ke[i * 24 + k], 1);
        if (rv != UO_OK) j->rv = rv;
      }
      if (s->threads > 1) pthread_mutex_unlock(&s->rowmtx[lf[i] - 1]);
    }
  }
}

static void *fill_entry(void *a) { fill_part((fill_job *)a); return NULL; }

/* fem_fill_K(AF_NO), fem/fem_sol.c:897-989: element blocks over threads */
int uo_fill_K(uo_sys *s, int npts)
{
  fill_job jobs[MAX_THREADS];
  pthread_t tid[MAX_THREADS];
  long quot, rem, i, n = s->m.n_elems;
  int rv = UO_OK;
  double t0 = now_s();
  if (s->threads <= 1) {
    jobs[0].s = s; jobs[0].from = 0; jobs[0].to = n; jobs[0].npts = npts;
    fill_part(&jobs[0]);
    rv = jobs[0].rv;
  } else {
    quot = n / s->threads; rem = n % s->threads;
    if (quot < 1) { quot = 0; rem = n; }
    for (i = 0; i < s->threads; i++) {
      jobs[i].s = s; jobs[i].npts = npts;
      jobs[i].from = i * quot; jobs[i].to = (i + 1) * quot;
      if (i == s->threads - 1) jobs[i].to += rem;
      pthread_create(&tid[i], NULL, fill_entry, &jobs[i]);
    }
    for (i = 0; i < s->threads; i++) pthread_join(tid[i], NULL);
    for (i = 0; i < s->threads; i++) if (jobs[i].rv != UO_OK) rv = jobs[i].rv;
  }
  s->t[1] = now_s() - t0;
  return rv;
}

/* fem_add_loads(0) -> femApplyNLoad type 3, fem/fem_sol.c:1119-1189,
 * fem/fem_loads.c:343-397 */
int uo_add_loads(uo_sys *s)
{
  long i;
  for (i = 0; i < s->m.n_nl; i++) {
    if (s->m.nl_type[i] != 3) continue;
    {
      long p = s->dofmap[s->m.nl_node[i] * KNOWN_DOFS + s->m.nl_dir[i] - 1];
      if (p < 1) return UO_ERR_VAL;
      s->F[p - 1] += s->m.nl_val[i];
    }
  }
  return UO_OK;
}

/* fem_add_disps(AF_YES, 0) -> femApplyNBC -> femApplyDisp with Val = 0,
 * fem/fem_sol.c:1195-1247, fem/fem_loads.c:116-208 */
int uo_add_disps(uo_sys *s)
{
  long i, r;
  double t0 = now_s();
  for (i = 0; i < s->m.n_nl; i++) {
    long p;
    double Fval, Val = s->m.nl_val[i];
    if (s->m.nl_type[i] != 1) continue;
    p = s->dofmap[s->m.nl_node[i] * KNOWN_DOFS + s->m.nl_dir[i] - 1];
    if (p < 1) return UO_ERR_VAL;
    if (Val != 0.0) {
      /* femApplyDisp, Val != 0 (fem/fem_loads.c:210-264): F -= K(:, p) Val
       * with the current K, row and column zeroed, u_p = Val and the
       * diagonal F_p / Val (or 1 with F_p = Val); no reaction bookkeeping */
      long k;
      for (k = 1; k <= s->K.rows; k++) s->F[k - 1] += -1.0 * rs_get(&s->K, k, p) * Val;
      rs_zero_row(&s->K, p);
      rs_zero_col(&s->K, p);
      s->u[p - 1] = Val;
      if (fabs(s->F[p - 1] / Val) > FEM_ZERO) {
        if (rs_putadd(&s->K, p, p, s->F[p - 1] / Val, 0) != UO_OK) return UO_ERR_BIG;
      } else {
        if (rs_putadd(&s->K, p, p, 1.0, 0) != UO_OK) return UO_ERR_BIG;
        s->F[p - 1] = Val;
      }
      continue;
    }
    Fval = s->F[p - 1];
    rs_zero_row(&s->K, p);
    rs_zero_col(&s->K, p);
    if (rs_putadd(&s->K, p, p, 1, 0) != UO_OK) return UO_ERR_BIG;
    s->u[p - 1] = 0;
    s->F[p - 1] = 0;
    for (r = 0; r < s->nreact; r++)              /* fem2IntPos, fem/fem_accs.c:495 */
      if (s->rnode[r] == s->m.nl_node[i] && s->rdof[r] == s->m.nl_dir[i]) break;
    if (r < s->nreact) s->rval[r] -= Fval;
  }
  s->t[2] = now_s() - t0;
  return UO_OK;
}

/* ---------------------------------------------------------------------- */
/* solvers                                                                  */
/* ---------------------------------------------------------------------- */

/* femMatNormBig, fem/fem_math.c:1856-1949 */
double uo_mat_norm(const uo_sys *s)
{ return fork_join(s->threads, OP_MATNORM, s->K.rows, &s->K, NULL, NULL, NULL); }

void uo_spmv(const uo_sys *s, const double *x, double *y) { spmv(s, x, y); }

/* femEqsCGwJ, fem/fem_eqs.c:88-247 */
static int cg_jacobi(uo_sys *s, double eps, long maxIt, int rule, long *iters, double *relres)
{
  long n = s->ndof, i, j;
  double *M, *r, *z, *p, *q, *x = s->u, *b = s->F;
  double ro, alpha, beta, roro = 0.0, normRes = 0.0, normX, normA, normB;
  int rv = UO_OK, converged = 0;
  *iters = 0;
  normA = uo_mat_norm(s);
  normB = vnorm(s, b, n);
  if (normB <= 0.0) {
    for (i = 0; i < n; i++) x[i] = 0.0;
    *relres = 0.0;
    return UO_OK;
  }
  M = (double *)calloc((size_t)n, sizeof(double));
  r = (double *)calloc((size_t)n, sizeof(double));
  z = (double *)calloc((size_t)n, sizeof(double));
  p = (double *)calloc((size_t)n, sizeof(double));
  q = (double *)calloc((size_t)n, sizeof(double));
  if (!M || !r || !z || !p || !q) { rv = UO_ERR_MEM; goto done; }
  for (i = 1; i <= n; i++) {
    M[i - 1] = rs_get(&s->K, i, i);
    if (fabs(M[i - 1]) < FEM_ZERO) { rv = UO_ERR_ZER; goto done; }
  }
  spmv(s, x, r);
  for (i = 0; i < n; i++) r[i] = b[i] - r[i];
  for (i = 1; i <= maxIt; i++) {
    for (j = 0; j < n; j++) z[j] = (r[j] / M[j]);
    ro = dot(s, r, z, n);
    if (i == 1) {
      for (j = 0; j < n; j++) p[j] = z[j];
    } else {
      beta = ro / roro;
      for (j = 0; j < n; j++) p[j] = (z[j] + (beta * p[j]));
    }
    spmv(s, p, q);
    alpha = ro / dot(s, p, q, n);
    for (j = 0; j < n; j++) {
      x[j] = x[j] + (alpha * p[j]);
      r[j] = r[j] - (alpha * q[j]);
    }
    normRes = vnorm(s, r, n);
    normX = vnorm(s, x, n);
    *iters = i;
    if (rule == 0) {
      if (normRes <= (eps * ((normA * normX) + normB))) { converged = 1; break; }
    } else {
      if (normRes <= eps * normB) { converged = 1; break; }
    }
    roro = ro;
  }
  if (!converged) rv = UO_ERR;
done:
  *relres = normRes / normB;
  free(M); free(r); free(z); free(p); free(q);
  return rv;
}

/* femEqsBiCCSwJ, fem/fem_eqs.c:262-468.  variant 1 keeps the reference's
 * convergence check eqsCompResid(a, b, x, &res) whose arguments are swapped
 * (it computes res = A b - x, fem/fem_eqs.c:37-75, :425) and, like the
 * reference, reports success whatever happens (:467).  Variant 2 checks the
 * true residual b - A x and reports non-convergence. */
static int bicgs_jacobi(uo_sys *s, int variant, double eps, long maxIt, int rule,
                        long *iters, double *relres)
{
  long n = s->ndof, i, j;
  double *M, *r, *rr, *p, *pp, *sv, *ss, *t, *v, *res, *x = s->u, *b = s->F;
  double ro, beta, roro = 0.0, alpha = 0.0, omega = 0.0, normRes = 0.0, normX, normA, normB;
  int rv = UO_OK, converged = 0;
  *iters = 0;
  normA = uo_mat_norm(s);
  normX = vnorm(s, x, n);
  normB = vnorm(s, b, n);
  M = (double *)calloc((size_t)n, sizeof(double));
  r = (double *)calloc((size_t)n, sizeof(double));
  rr = (double *)calloc((size_t)n, sizeof(double));
  p = (double *)calloc((size_t)n, sizeof(double));
  pp = (double *)calloc((size_t)n, sizeof(double));
  sv = (double *)calloc((size_t)n, sizeof(double));
  ss = (double *)calloc((size_t)n, sizeof(double));
  t = (double *)calloc((size_t)n, sizeof(double));
  v = (double *)calloc((size_t)n, sizeof(double));
  res = (double *)calloc((size_t)n, sizeof(double));
  if (!M || !r || !rr || !p || !pp || !sv || !ss || !t || !v || !res) { rv = UO_ERR_MEM; goto done; }
  for (i = 1; i <= n; i++) {
    M[i - 1] = rs_get(&s->K, i, i);
    if (fabs(M[i - 1]) < FEM_ZERO) { rv = UO_ERR_ZER; goto done; }
  }
  spmv(s, x, r);
  for (i = 0; i < n; i++) { r[i] = b[i] - r[i]; rr[i] = r[i]; }
  if (vnorm(s, r, n) <= FEM_ZERO) { converged = 1; goto done; }
  for (i = 1; i <= maxIt; i++) {
    ro = dot(s, rr, r, n);
    if (fabs(ro) <= 0.0) break;
    if (i == 1) {
      for (j = 0; j < n; j++) p[j] = r[j];
    } else {
      beta = (ro / roro) * (alpha / omega);
      for (j = 0; j < n; j++) p[j] = r[j] + beta * (p[j] - (omega * v[j]));
    }
    for (j = 0; j < n; j++) pp[j] = p[j] / M[j];
    spmv(s, pp, v);
    alpha = (ro / dot(s, rr, v, n));
    for (j = 0; j < n; j++) sv[j] = r[j] - (alpha * v[j]);
    *iters = i;
    if (vnorm(s, sv, n) <= FEM_ZERO) {
      for (j = 0; j < n; j++) x[j] += alpha * pp[j];
      converged = 1;
      break;
    }
    for (j = 0; j < n; j++) ss[j] = sv[j] / M[j];
    spmv(s, ss, t);
    omega = (dot(s, t, sv, n) / dot(s, t, t, n));
    for (j = 0; j < n; j++) {
      x[j] += alpha * pp[j] + (omega * ss[j]);
      r[j] = sv[j] - (omega * t[j]);
    }
    roro = ro;
    if (variant == 1) {          /* eqsCompResid(a, b, x, &res): res = A b - x,
                                    accumulated onto -x, fem/fem_eqs.c:48-58 */
      const rs_mat *A = &s->K;
      long ii, jj;
      for (ii = 0; ii < n; ii++) {
        res[ii] = 0 - x[ii];
        for (jj = A->frompos[ii]; jj < A->frompos[ii] + A->defpos[ii]; jj++) {
          if (A->pos[jj] <= 0) break;
          res[ii] += (A->data[jj] * b[A->pos[jj] - 1]);
        }
      }
    } else {                     /* true residual b - A x */
      spmv(s, x, res);
      for (j = 0; j < n; j++) res[j] = b[j] - res[j];
    }
    normRes = vnorm(s, res, n);
    normX = vnorm(s, x, n);
    if (rule == 0) {
      if (normRes < (eps * ((normA * normX) + normB))) { converged = 1; break; }
    } else {
      if (normRes < eps * normB) { converged = 1; break; }
    }
  }
done:
  *relres = normB > 0 ? normRes / normB : 0.0;
  free(M); free(r); free(rr); free(p); free(pp); free(sv); free(ss); free(t); free(v); free(res);
  if (rv != UO_OK) return rv;
  if (variant == 1) return UO_OK;
  return converged ? UO_OK : UO_ERR;
}

/* femEqsCGwSSOR, fem/fem_eqs.c:999-1188, the MAT_SPAR branch (:1086-1112):
 * symmetric Gauss-Seidel in natural order, zz = M (r - L zz) forward over
 * the slots with 1 <= column < row, then z = zz - M U z backward over the
 * slots with column > row (M = 1 / diag). */
static int cg_ssor(uo_sys *s, double eps, long maxIt, int rule, long *iters, double *relres)
{
  const rs_mat *a = &s->K;
  long n = s->ndof, i, ii, j, ipos;
  double *M, *r, *z, *zz, *p, *q, *x = s->u, *b = s->F;
  double ro, alpha, beta, roro = 0.0, normRes = 0.0, normX, normA, normB, val;
  int rv = UO_OK, converged = 0;
  *iters = 0;
  normA = uo_mat_norm(s);
  normB = vnorm(s, b, n);
  if (normB <= 0.0) {
    for (i = 0; i < n; i++) x[i] = 0.0;
    *relres = 0.0;
    return UO_OK;
  }
  M = (double *)calloc((size_t)n, sizeof(double));
  r = (double *)calloc((size_t)n, sizeof(double));
  z = (double *)calloc((size_t)n, sizeof(double));
  zz = (double *)calloc((size_t)n, sizeof(double));
  p = (double *)calloc((size_t)n, sizeof(double));
  q = (double *)calloc((size_t)n, sizeof(double));
  if (!M || !r || !z || !zz || !p || !q) { rv = UO_ERR_MEM; goto done; }
  for (i = 1; i <= n; i++) {
    val = rs_get(a, i, i);
    if (fabs(val) < FEM_ZERO) { rv = UO_ERR_ZER; goto done; }
    M[i - 1] = 1.0 / val;
  }
  spmv(s, x, r);
  for (i = 0; i < n; i++) r[i] = b[i] - r[i];
  for (i = 1; i <= maxIt; i++) {
    for (ii = 1; ii <= n; ii++) {
      val = 0.0;
      for (j = a->frompos[ii - 1]; j < a->frompos[ii - 1] + a->defpos[ii - 1]; j++) {
        ipos = a->pos[j];
        if (ipos >= ii || ipos < 1) continue;
        val += a->data[j] * zz[ipos - 1];
      }
      zz[ii - 1] = M[ii - 1] * (r[ii - 1] - val);
    }
    for (ii = n; ii >= 1; ii--) {
      val = 0.0;
      for (j = a->frompos[ii - 1]; j < a->frompos[ii - 1] + a->defpos[ii - 1]; j++) {
        ipos = a->pos[j];
        if (ipos > ii) val += a->data[j] * z[ipos - 1];
      }
      z[ii - 1] = zz[ii - 1] - M[ii - 1] * val;
    }
    ro = dot(s, r, z, n);
    if (i == 1) {
      for (j = 0; j < n; j++) p[j] = z[j];
    } else {
      beta = ro / roro;
      for (j = 0; j < n; j++) p[j] = (z[j] + (beta * p[j]));
    }
    spmv(s, p, q);
    alpha = ro / dot(s, p, q, n);
    for (j = 0; j < n; j++) {
      x[j] = x[j] + (alpha * p[j]);
      r[j] = r[j] - (alpha * q[j]);
    }
    normRes = vnorm(s, r, n);
    normX = vnorm(s, x, n);
    *iters = i;
    if (rule == 0) {
      if (normRes <= (eps * ((normA * normX) + normB))) { converged = 1; break; }
    } else {
      if (normRes <= eps * normB) { converged = 1; break; }
    }
    roro = ro;
  }
  if (!converged) rv = UO_ERR;
done:
  *relres = normRes / normB;
  free(M); free(r); free(z); free(zz); free(p); free(q);
  return rv;
}

int uo_solve(uo_sys *s, int solver, double eps, long maxit, int rule, long *iters, double *relres)
{
  int rv;
  double t0 = now_s();
  if (maxit < 0) maxit = s->ndof;
  if (solver == 0) rv = cg_jacobi(s, eps, maxit, rule, iters, relres);
  else if (solver == 3) rv = cg_ssor(s, eps, maxit, rule, iters, relres);
  else rv = bicgs_jacobi(s, solver, eps, maxit, rule, iters, relres);
  s->t[3] = now_s() - t0;
  return rv;
}

/* fem_fill_K(AF_YES) restricted to what the reactions need:
 * femLocUtoU_e (fem/fem_mloc.c:301-341), element forces (see
 * element_forces), femLocF_e into F (fem/fem_mloc.c:230-293) and
 * fem_fill_K_React (fem/fem_sol.c:844-871).  Serial, element order. */
int uo_reactions(uo_sys *s, int npts)
{
  double xyz[24], ue[24], Fe[24];
  long lf[24], e, i;
  int k, rv;
  double t0 = now_s();
  for (e = 0; e < s->m.n_elems; e++) {
    elem_xyz(&s->m, e, xyz);
    elem_locvec(s, e, lf);
    for (k = 0; k < 24; k++) ue[k] = lf[k] ? 0.0 + s->u[lf[k] - 1] : 0.0 + 0.0;
    rv = element_forces(xyz, s->m.E[e], s->m.nu[e], npts, ue, Fe, NULL);
    if (rv != UO_OK) return rv;
    for (k = 0; k < 24; k++)
      if (lf[k]) s->F[lf[k] - 1] += Fe[k];
  }
  for (i = 0; i < s->nreact; i++) {
    long p = s->dofmap[s->rnode[i] * KNOWN_DOFS + s->rdof[i] - 1];
    if (p <= 0) return UO_ERR_VAL;
    s->rval[i] += s->F[p - 1];
  }
  s->t[4] = now_s() - t0;
  return UO_OK;
}

/* the element results of fem_fill_K(AF_YES) for the current u: res
 * [n_elems][npts^3][12] (see element_forces) */
int uo_element_results(const uo_sys *s, int npts, double *res)
{
  double xyz[24], ue[24], Fe[24];
  long lf[24], e;
  int k, rv;
  for (e = 0; e < s->m.n_elems; e++) {
    elem_xyz(&s->m, e, xyz);
    elem_locvec(s, e, lf);
    for (k = 0; k < 24; k++) ue[k] = lf[k] ? 0.0 + s->u[lf[k] - 1] : 0.0 + 0.0;
    rv = element_forces(xyz, s->m.E[e], s->m.nu[e], npts, ue, Fe, res + 12L * npts * npts * npts * e);
    if (rv != UO_OK) return rv;
  }
  return UO_OK;
}

long uo_K_nnz(const uo_sys *s)
{
  long i, j, c = 0;
  for (i = 0; i < s->K.rows; i++)
    for (j = s->K.frompos[i]; j < s->K.frompos[i] + s->K.defpos[i]; j++) {
      if (s->K.pos[j] <= 0) break;
      c++;
    }
  return c;
}

void uo_get_K(const uo_sys *s, long *rowlen, long *cols, double *vals)
{
  long i, j, c = 0;
  for (i = 0; i < s->K.rows; i++) {
    long r0 = c;
    for (j = s->K.frompos[i]; j < s->K.frompos[i] + s->K.defpos[i]; j++) {
      if (s->K.pos[j] <= 0) break;
      cols[c] = s->K.pos[j];
      vals[c] = s->K.data[j];
      c++;
    }
    rowlen[i] = c - r0;
  }
}

void uo_get_F(const uo_sys *s, double *F) { memcpy(F, s->F, sizeof(double) * (size_t)s->ndof); }
void uo_get_u(const uo_sys *s, double *u) { memcpy(u, s->u, sizeof(double) * (size_t)s->ndof); }
void uo_set_u(uo_sys *s, const double *u) { memcpy(s->u, u, sizeof(double) * (size_t)s->ndof); }
void uo_get_R(const uo_sys *s, double *R, long *rnode, long *rdof)
{
  long i;
  for (i = 0; i < s->nreact; i++) { R[i] = s->rval[i]; rnode[i] = s->rnode[i]; rdof[i] = s->rdof[i]; }
}
void uo_get_dofmap(const uo_sys *s, long *d)
{ memcpy(d, s->dofmap, sizeof(long) * (size_t)s->m.n_nodes * KNOWN_DOFS); }
void uo_timings(const uo_sys *s, double t[5]) { memcpy(t, s->t, sizeof(double) * 5); }
