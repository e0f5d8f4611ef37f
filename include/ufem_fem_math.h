/*
 * ufem_fem_math.h -- drop-in entry points of libufem_gpu.so that replace
 * uFEM's own symbols on the linear-static solve path.
 *
 * The matrix/vector types keep the reference layout exactly (same members,
 * same order, LP64 `long`), so a tMatrix built by uFEM's fem_math.c can be
 * passed straight in:  fem/fem_math.h:47-67.  Enum values: fem/fem_math.h:32-39.
 * Return codes: fem/fem.h:41-52.
 *
 * Every function below has the exact signature and argument meaning of the
 * reference function it replaces (cited per function).  Matrices are the
 * reference's row-slot sparse form (MAT_SPAR) with 1-based column ids, slots
 * filled in insertion order and 0 marking the first empty slot
 * (fem/fem_math.c:185-235).  The GPU serves MAT_SPAR matrices and VEC_FULL
 * vectors.  A call with any other type (a dense matrix, a sparse vector) is
 * handed to the reference's own CPU function when the program links it under
 * the name <function>_cpu (fem_math.o and fem_eqs.o compiled with
 * -Dfunction=function_cpu, see the integration notes); without it the call
 * returns AF_ERR_TYP (NAN for the functions returning double).
 *
 * The GPU keeps a device copy of every matrix it is handed (a side table
 * keyed by the data pointer, len, rows and cols; see ufem_gpu.h for the
 * policy calls ufem_gpu_set_mirror_policy, ufem_gpu_mark_dirty and
 * ufem_gpu_release).
 */
#ifndef UFEM_FEM_MATH_H
#define UFEM_FEM_MATH_H

#ifdef __cplusplus
extern "C" {
#endif

/* A source of the reference that has included its own fem_math.h (and with
 * it fem.h) already holds these definitions, with the same values and
 * members; they are skipped there so that it can include this header. */
#ifndef __FEM_MATH_H__
#define MAT_FULL 0
#define MAT_SPAR 1
#define VEC_FULL 0
#define VEC_SPAR 1

#define AF_OK       0
#define AF_ERR     -1
#define AF_ERR_IO  -2
#define AF_ERR_VAL -3
#define AF_ERR_MEM -4
#define AF_ERR_TYP -5
#define AF_ERR_SIZ -9
#define AF_ERR_BIG -11
#define AF_ERR_ZER -13

typedef struct
{
  long    type;     /* 0 = dense; 1 = sparse (row slots)              */
  long    rows;
  long    cols;
  long    len;      /* length of pos and data                         */
  long   *pos;      /* sparse: 1-based column, 0 = empty slot         */
  double *data;
  long   *frompos;  /* sparse: first slot of each row                 */
  long   *defpos;   /* sparse: slot capacity of each row              */
} tMatrix;

typedef struct
{
  long    type;
  long    rows;
  long    len;
  long   *pos;
  double *data;
} tVector;
#endif

/* Jacobi-preconditioned CG.  a, b read-only; x is the start vector in and the
 * solution out.  Stops when ||r|| <= eps (||A|| ||x|| + ||b||) with ||A|| the
 * largest row 2-norm (the reference rule), or -- after
 * ufem_gpu_set_stop_rule(1) -- when ||r|| <= eps ||b||.  Returns AF_OK,
 * AF_ERR if maxIt iterations do not converge, AF_ERR_ZER on a diagonal
 * smaller than 1e-17, AF_ERR_SIZ on mismatched sizes.
 * Replaces femEqsCGwJ, fem/fem_eqs.c:88-247 (declared fem/fem_math.h:153). */
int femEqsCGwJ(tMatrix *a, tVector *b, tVector *x, double eps, long maxIt);

/* Jacobi (right-) preconditioned BiCGStab, same arguments.  Convergence is
 * tested on the true residual b - A x, which is what the reference documents;
 * its own test passes (b, x) swapped to eqsCompResid and effectively stops
 * only when ||s|| <= 1e-17 (fem/fem_eqs.c:425).  As the reference does
 * (fem/fem_eqs.c:441-467), a run that ends unconverged (maxIt reached, or a
 * stop the true residual did not confirm after the restarts) keeps its last
 * x and returns AF_OK, with a line on stderr; ufem_gpu_last_solve() then
 * reports status AF_ERR.  A diagonal smaller than 1e-17 returns AF_ERR_ZER
 * (the reference jumps to its exit there and returns AF_OK, x untouched).
 * Replaces femEqsBiCCSwJ, fem/fem_eqs.c:262-468 (fem/fem_math.h:155). */
int femEqsBiCCSwJ(tMatrix *a, tVector *b, tVector *x, double eps, long maxIt);

/* CG with the SSOR (omega = 1, symmetric Gauss-Seidel) preconditioner: the
 * same arguments, stopping rule and return codes as femEqsCGwJ.  The sweeps
 * run in a greedy colour order of the rows instead of the reference's
 * natural order, so the preconditioner (and the iteration count) differ
 * while x solves the same system.  Replaces femEqsCGwSSOR,
 * fem/fem_eqs.c:999-1188 (fem/fem_math.h:154). */
int femEqsCGwSSOR(tMatrix *a, tVector *b, tVector *x, double eps, long maxIt);

/* c = a b.  Replaces femMatVecMultBig, fem/fem_math.c:1350-1451. */
int femMatVecMultBig(tMatrix *a, tVector *b, tVector *c);

/* a . b.  Replaces femVecVecMultBig, fem/fem_math.c:1017-1100. */
double femVecVecMultBig(tVector *a, tVector *b);

/* ||a||_2.  Replaces femVecNormBig, fem/fem_math.c:2022-2100. */
double femVecNormBig(tVector *a);

/* max_i ||A_i,:||_2.  Replaces femMatNormBig, fem/fem_math.c:1856-1949. */
double femMatNormBig(tMatrix *a);

/* c = amult a + bmult b.  Replaces femVecLinCombBig, fem/fem_math.c:1543. */
int femVecLinCombBig(double amult, tVector *a, double bmult, tVector *b, tVector *c);

/* every stored value (a->len of them) to 0, the slots' columns kept; a device
 * copy of the matrix is cleared with it.  Replaces femMatSetZeroBig,
 * fem/fem_math.c:483-523 (fem/fem_math.h:107). */
void femMatSetZeroBig(tMatrix *a);

/* every stored value (a->len) to 0.  Replaces femVecSetZeroBig,
 * fem/fem_math.c:920-958 (fem/fem_math.h:125). */
void femVecSetZeroBig(tVector *a);


#ifdef __cplusplus
}
#endif
#endif
