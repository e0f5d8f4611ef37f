/*
 * ufem_oracle.h -- clean-room CPU restatement of uFEM 0.5.3's linear-static
 * solve path.  TEST INFRASTRUCTURE ONLY: the parity tests, the smoke check and
 * the CPU-baseline leg of the benchmark load it as the checker; the product
 * library never links or calls it.
 *
 * Functions restated (each cites the reference file:line in ufem_oracle.c):
 *   element 9 stiffness (27 or 8 Gauss points)   fem/fem_e009.c:261-635
 *   DOF numbering and row capacities             fem/fem_sol.c:1263-1391, fem/fem_mloc.c:83-116
 *   row-slot sparse matrix                       fem/fem_math.c:57-235
 *   assembly in element order                    fem/fem_mloc.c:163-221
 *   nodal loads / zero-displacement supports     fem/fem_loads.c:116-208, 343-397
 *   Jacobi-PCG and Jacobi-BiCGStab               fem/fem_eqs.c:88-247, 262-468
 *   reactions (results pass)                     fem/fem_sol.c:844-871, fem/fem_e009.c:417-580
 *   pthread fork-join kernels (<= 8 threads)     fem/fem_math.c:988-1100, 1321-1451, 1819-2100
 */
#ifndef UFEM_ORACLE_H
#define UFEM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* mesh + loads in the reference's preparsed-input terms (fem/feminput.5) */
typedef struct {
  long n_nodes;
  const double *x, *y, *z;     /* node coordinates                       */
  long n_elems;
  const long *enodes;          /* 8 node positions (0-based) per brick   */
  const double *E, *nu;        /* per-element material-1 values          */
  long n_nl;                   /* nodal loads and supports, input order  */
  const long *nl_node;         /* node position (0-based)                */
  const long *nl_type;         /* 1 = displacement, 3 = force            */
  const long *nl_dir;          /* 1 = ux, 2 = uy, 3 = uz                 */
  const double *nl_val;
} uo_model;

typedef struct uo_sys uo_sys;

uo_sys *uo_create(const uo_model *m, int threads);
void    uo_free(uo_sys *s);
long    uo_ndof(const uo_sys *s);
long    uo_nreact(const uo_sys *s);

/* element stiffness of one brick, row-major 24x24; npts = 2 or 3 */
int  uo_element_stiffness(const double xyz[24], double E, double nu, int npts, double Ke[576]);

int  uo_fill_K(uo_sys *s, int npts);            /* fem_fill_K(AF_NO)            */
int  uo_add_loads(uo_sys *s);                   /* fem_add_loads(0)             */
int  uo_add_disps(uo_sys *s);                   /* fem_add_disps(AF_YES, 0)     */
/* solver: 0 = femEqsCGwJ, 1 = femEqsBiCCSwJ (as written), 2 = BiCGStab with
 * the true residual b - A x, 3 = femEqsCGwSSOR.  rule: 0 = reference criterion
 * ||r|| <= eps (||A|| ||x|| + ||b||); 1 = ||r|| <= eps ||b||. */
int  uo_solve(uo_sys *s, int solver, double eps, long maxit, int rule,
              long *iters, double *relres);
int  uo_reactions(uo_sys *s, int npts);         /* fem_fill_K(AF_YES) + reactions */
/* element results of fem_fill_K(AF_YES) for the current u:
 * res [n_elems][npts^3][12] = sigma x..xy, epsilon x..xy at each point */
int  uo_element_results(const uo_sys *s, int npts, double *res);

/* accessors: copy out state (arrays sized by the caller) */
long uo_K_nnz(const uo_sys *s);
void uo_get_K(const uo_sys *s, long *rowlen, long *cols, double *vals); /* slot order, 1-based cols */
void uo_get_F(const uo_sys *s, double *F);
void uo_get_u(const uo_sys *s, double *u);
void uo_set_u(uo_sys *s, const double *u);
void uo_get_R(const uo_sys *s, double *R, long *rnode, long *rdof);
void uo_get_dofmap(const uo_sys *s, long *nDOFfld);  /* 7 per node, 0 = inactive */
double uo_mat_norm(const uo_sys *s);           /* femMatNormBig               */
void uo_spmv(const uo_sys *s, const double *x, double *y); /* femMatVecMultBig  */

/* wall seconds of the phases of the last calls */
void uo_timings(const uo_sys *s, double t[5]); /* dofs, fill, bc, solve, react */

#ifdef __cplusplus
}
#endif
#endif
