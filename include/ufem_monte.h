/*
 * ufem_monte.h -- the Monte plugin interface of libufem_gpu.so: the seven
 * functions uFEM exports from its libfem.so for the Monte probabilistic
 * program (fem/fem_mont.c built with USE_MONTE, fem/Makefile:192-193), same
 * names, signatures and argument meaning, with the linear-static solution of
 * every sample on the GPU.
 *
 * The model is a preparsed uFEM input file (fem/feminput.5) of 8-node bricks
 * (element type 9), material type 1, nodal forces and supports, ending in the
 * random-variable section read by fem_monte_read_data (fem/fem_mont.c:866-909):
 *   "ilen olen", then ilen + olen rows "pos type item subitem subitemiter"
 * with the variable codes of fem/fem_mont.h:30-50 and 0-based item positions.
 * Inputs acted on: MAT (1, a value of a material set: EX and NU change the
 * stiffness), N (2, a node coordinate), NLD (3) and NLPOS (5) (a nodal load
 * or support value), MAX_D (7).  RS, DAMP and EL are accepted and change
 * nothing on this path (no real-set values for the brick, no dynamics, no
 * element loads).  Outputs: RES_D, RES_R, RES_E, RES_SUM_E, RES_MAX_E,
 * RES_MIN_E, RES_FAIL_E, RES_D_MAX, RES_D_MIN, RES_D_SUM, RES_A_MAX, RES_FAIL,
 * computed as monte_fill_ofld_data does (fem/fem_mont.c:514-691).
 *
 * One model per process (the reference keeps it in globals too).  Error text:
 * ufem_gpu_last_error() (include/ufem_gpu.h).
 */
#ifndef UFEM_MONTE_H
#define UFEM_MONTE_H

#ifdef __cplusplus
extern "C" {
#endif

/* 2, the "advanced" interface (fem/fem_mont.c:137-141) */
long monte_dlib_interface_type(void);
/* numbers of input and output variables and the output index of the failure
 * function, -1 if none (fem/fem_mont.c:146-153) */
void monte_nums_of_vars(char *param, long *ilen, long *olen, long *ffunc);
/* names of variable `pos` (0..n-1), "empty" past the end; the string is
 * static and overwritten by the next call (fem/fem_mont.c:158-250, :256-398) */
char *monte_ivar_name(char *param, long pos);
char *monte_ovar_name(char *param, long pos);
/* reads the input file named by `param`; 0 or -1 (fem/fem_mont.c:424-469) */
int monte_init_lib_stuff(char *param);
/* frees the model and the device system; 0 (fem/fem_mont.c:471-477) */
int monte_clean_lib_stuff(char *param);
/* one sample: applies ifld, solves (femEqsCGwSSOR with eps 1e-21 and 3 n
 * iterations at most, as monte_linear_solver, fem/fem_mont.c:480-511) and
 * fills ofld; returns AF_OK or the solver's AF_* code (ofld then stays 0).
 * if_type 1 (price) and 2 (dynamics) return AF_ERR_TYP
 * (fem/fem_mont.c:694-855) */
int monte_solution(char *param, double *ifld, double *ofld, long if_type);

/* No counterpart in the reference: n samples as slots of one batched GPU
 * launch, for a Monte driver that can hand over several samples at once.
 * Sample k reads ifld[k ilen .. k ilen + ilen - 1] and writes
 * ofld[k olen .. k olen + olen - 1] (every variable's slot must lie inside
 * its field, else AF_ERR_SIZ); the outputs equal those of n monte_solution
 * calls in sample order (reaction sums included), with each sample solved by
 * Jacobi PCG to ||r|| <= 1e-15 ||f|| instead of SSOR-CG.  A model with a
 * support of non-zero value, or inputs that set one, is solved one sample at
 * a time.  status[k] (may be NULL): AF_OK or AF_ERR (not converged) per
 * sample; returns the first failing code. */
int monte_solution_batch(char *param, long n, const double *ifld, double *ofld, long if_type, int *status);

#ifdef __cplusplus
}
#endif
#endif
