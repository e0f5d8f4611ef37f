/*
 * ufem_gpu.h -- device-resident linear-static solve of libufem_gpu.so.
 *
 * These entry points replace, for meshes of 8-node bricks (element type 9)
 * with isotropic material (type 1), the steps femSolve runs on the CPU
 * (fem/fem_sol.c:1655-1817):
 *
 *   ufem_sys_create      fem_dofs + femKhit + fem_sol_alloc + fem_sol_res_alloc
 *                        (fem/fem_sol.c:1263-1391, :485, :639; fem/fem_mloc.c:83)
 *   ufem_sys_fill_K      fem_fill_K(AF_NO)       (fem/fem_sol.c:897-989)
 *   ufem_sys_add_loads   fem_add_loads(0)        (fem/fem_sol.c:1119-1189)
 *   ufem_sys_add_disps   fem_add_disps(AF_YES,0) (fem/fem_sol.c:1195-1247)
 *   ufem_sys_solve       femEqsCGwJ / femEqsBiCCSwJ / femEqsCGwSSOR (fem/fem_eqs.c:88, :262, :999)
 *   ufem_sys_reactions   fem_fill_K(AF_YES) + fem_fill_K_React (fem/fem_sol.c:844)
 *
 * The mesh arrives in the reference's preparsed-input terms (fem/feminput.5):
 * node coordinates, 8 node positions per brick in the reference node order
 * (fem/fem_e009.c:36-69), per-element E and nu, and the nodal load/support
 * list in input order (type 1 = prescribed displacement, any value;
 * type 3 = force; direction 1..3 = ux, uy, uz).  Every node must belong to an
 * element; DOF 3 p + (dir - 1) then equals the reference's node-major
 * numbering.  Return codes are the reference's AF_* values.
 */
#ifndef UFEM_GPU_H
#define UFEM_GPU_H

#include "ufem_fem_math.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  long n_nodes;
  const double *x, *y, *z;
  long n_elems;
  const long *enodes;          /* 8 per element, 0-based node positions   */
  const double *E, *nu;        /* per element                            */
  long n_nl;
  const long *nl_node;         /* 0-based node position                  */
  const long *nl_type;         /* 1 displacement, 3 force                */
  const long *nl_dir;          /* 1 ux, 2 uy, 3 uz                       */
  const double *nl_val;
} ufem_mesh;

typedef struct ufem_sys ufem_sys;

enum {
  UFEM_SOLVER_PCG = 0,         /* Jacobi PCG, FP64                        */
  UFEM_SOLVER_BICGSTAB = 1,    /* Jacobi BiCGStab, FP64                   */
  UFEM_SOLVER_MIXED = 2,       /* FP32 inner PCG in FP64 refinement       */
  UFEM_SOLVER_LEARNED = 3,     /* flexible CG with the learned preconditioner */
  UFEM_SOLVER_LEARNED_MIXED = 4, /* the same with q = A p from the FP32 copy of A and
                                   FP64 residual refreshes (inner_eps, inner_maxit) */
  UFEM_SOLVER_SSOR = 5,         /* CG with the SSOR (symmetric Gauss-Seidel) preconditioner,
                                   rows in greedy colour order */
  UFEM_SOLVER_TWOLEVEL = 6,     /* PCG with z = dinv r + P A^-1 P^T r, A = P^T K P on a coarse
                                   node grid (ufem_sys_set_coarse first) */
  UFEM_SOLVER_TWOLEVEL_MIXED = 7, /* the same with q = K p from the FP32 copy of K and FP64
                                   residual refreshes (inner_eps, inner_maxit), convergence
                                   accepted on a refreshed residual only */
  UFEM_SOLVER_MIXED_V64 = 8     /* Jacobi PCG with q = K p from the FP32 copy of K and every
                                   vector, dot and recurrence in FP64; FP64 residual
                                   refreshes (inner_eps, inner_maxit), convergence accepted
                                   on a refreshed residual only */
};
enum { UFEM_RULE_REFERENCE = 0, UFEM_RULE_RELATIVE = 1 };
/* streamed: single-reduction PCG, two kernels per iteration (w = K u with
 * (w,u); one vector kernel with the reduction, test and all updates),
 * replayed from a HIP graph;  persistent: one cooperative launch with grid
 * barriers;  fused: classic PCG, two kernels, the p update folded into the
 * SpMV gather;  classic: classic PCG, three kernels (p update, SpMV + (p,q),
 * vector update) */
enum { UFEM_EXEC_STREAMED = 0, UFEM_EXEC_PERSISTENT = 1, UFEM_EXEC_FUSED = 2, UFEM_EXEC_CLASSIC = 3 };

typedef struct {
  int solver;
  int rule;          /* 0: ||r|| <= eps (||A|| ||x|| + ||b||); 1: ||r|| <= eps ||b|| */
  double eps;
  long maxit;        /* 0 means n, as femSolve passes nDOFAct; < 0: none  */
  int exec;          /* streamed kernels (graph replayed) or one persistent launch */
  int chunk;         /* streamed form: iterations per graph replay (0: 32) */
  double inner_eps;  /* mixed: the FP64 residual is refreshed each time the FP32
                        residual has dropped by this factor (0: 0.1) */
  long inner_maxit;  /* mixed: ... and at least every this many iterations (0: 500) */
  void *net;         /* learned: network handle from ufem_net_*           */
  int warm_start;    /* nonzero: start from the current u                 */
  int fcg_m;         /* learned: directions kept by flexible CG (1..8; 0: 1) */
  int coarse;        /* learned, learned-mixed: add the two-level correction
                        (ufem_sys_set_coarse first); net may then be NULL,
                        which leaves dinv r + the correction */
  int inner_restart; /* mixed: classical iterative refinement -- every FP64
                        refresh starts a new FP32 CG on the correction
                        equation, run until its residual has fallen by
                        inner_eps (or inner_maxit iterations); 0: the FP32 CG
                        keeps its direction across refreshes (reliable update) */
  int f32_rem;       /* mixed forms with FP64 vectors (learned-mixed,
                        two-level-mixed, mixed-v64): the FP32 copy of K also
                        keeps each value's remainder K - fl32(K) rounded to
                        bf16 (48-bit values, 6 bytes instead of 8); FP32
                        vectors (UFEM_SOLVER_MIXED) ignore it */
  int net_prec;      /* learned: how the network is applied (coarse-first
                        form, 16 or 32 channels): 0 FP64 activations, or,
                        with the two-level correction, f32; 1 FP64; 2 f32
                        activations on the bf16 matrix cores (the training
                        path).  Flexible CG loses iterations to f32 rounding
                        noise in the preconditioner, which FP64 removes at a
                        higher cost per iteration; with the coarse term the
                        solve is short and the noise costs little */
} ufem_solve_opts;

typedef struct {
  int status;        /* AF_OK, AF_ERR (not converged), AF_ERR_ZER, ...    */
  long iters;        /* iterations (all inner iterations for mixed)       */
  long outer;        /* mixed, learned-mixed: FP64 residual refreshes;
                        learned: iteration of the Jacobi fallback;
                        SSOR: colours of the sweep order;
                        BiCGStab: restarts after a stop that the FP64
                        true residual did not confirm                     */
  int fallback;      /* learned: 1 if Jacobi-PCG had to take over         */
  double relres;     /* last ||r||/||b|| the solver tested                */
  double true_relres;/* ||b - K u|| / ||b|| recomputed in FP64            */
  double solve_ms;   /* device time of the solve (HIP events)             */
} ufem_solve_info;

ufem_sys *ufem_sys_create(const ufem_mesh *m);
void ufem_sys_destroy(ufem_sys *s);
long ufem_sys_ndof(const ufem_sys *s);
long ufem_sys_nnz(const ufem_sys *s);       /* scalar entries of K, explicit zeros kept */
long ufem_sys_ncolors(const ufem_sys *s);
long ufem_sys_nreact(const ufem_sys *s);
/* new Young's moduli, one per element, on the same mesh (Monte-Carlo and
 * training fields); fill_K and add_disps must follow */
int  ufem_sys_set_E(ufem_sys *s, const double *E);
/* new node coordinates, E, nu and load/support values on the same mesh
 * (same sizes, connectivity and load/support list), and F, K and the
 * reaction slots back to the state a solve starts from -- what
 * monte_linear_solver does between samples (fem/fem_mont.c:466-487);
 * fill_K, add_loads and add_disps must follow */
int  ufem_sys_update(ufem_sys *s, const ufem_mesh *m);
int  ufem_sys_fill_K(ufem_sys *s, int npts);   /* npts = 2 (exact for parallelepipeds) or 3 */
int  ufem_sys_add_loads(ufem_sys *s);
int  ufem_sys_add_disps(ufem_sys *s);
int  ufem_sys_solve(ufem_sys *s, const ufem_solve_opts *o, ufem_solve_info *info);
/* No counterpart in the reference.  For callers that cannot allocate (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc):
 * the solve of ufem_sys_solve for o->solver = UFEM_SOLVER_PCG, exec
 * UFEM_EXEC_STREAMED, calling no hipMalloc, hipFree, pinned-memory, event or
 * graph routine.  Its vectors, partial sums and scalars live in one
 * file-scope device array per GPU (about 52 MiB, held until the process
 * ends) that bounds the system to 1,048,576 unknowns (AF_ERR_BIG beyond;
 * AF_ERR_TYP for another solver or form).  All calls on one GPU share that
 * array: calls that overlap on one GPU race.  The result equals
 * ufem_sys_solve's; solve_ms is host wall time. */
int  ufem_sys_solve_noalloc(ufem_sys *s, const ufem_solve_opts *o, ufem_solve_info *info);
/* reactions, one per support entry in input order (node position, direction,
 * value); supports of non-zero prescribed value are included here, though the
 * reference's own list leaves them out (fem/fem_sol.c:676-722) */
int  ufem_sys_reactions(ufem_sys *s, double *R, long *rnode, long *rdof);
/* strains and stresses at the npts^3 Gauss points (npts 2 or 3) of every
 * element for the current u, the results pass of fem_fill_K(AF_YES)
 * (fem/fem_e009.c:417-535, material type 1): res [n_elems][npts^3][12] =
 * sigma x, y, z, yz, zx, xy then epsilon x, y, z, yz, zx, xy (engineering
 * shears), point (i npts + j) npts + k along the local axes */
int  ufem_sys_element_results(ufem_sys *s, int npts, double *res);
/* average device time (ms) of the results pass, `reps` launches into a
 * device buffer; negative on error */
double ufem_sys_time_results(ufem_sys *s, int npts, int reps);
/* the reference's result file (femWriteRes, fem/fem_io.c:648-734) for the
 * current u, with the 27-point element results in the reference's element
 * result layout; sol_id, dep_on and sim_num fill the header line (the input
 * file's solution ID and dependency, the simulation number); node and
 * element IDs are written as 1-based positions; the reactions of supports of
 * zero value only, as the reference lists them */
int  ufem_sys_write_results(ufem_sys *s, const char *fname, long sol_id, long dep_on, double sim_num);
/* node and element IDs the result file names (the input's IDs, see
 * ufem_model_*); NULL: 1-based positions */
int  ufem_sys_set_ids(ufem_sys *s, const long *node_id, const long *elem_id);
/* the result file of the fem binary's -fbc (fem/fem_sol.c:1280-1296): DOFs
 * of zero-valued supports left out of the DOF map and of u, their reactions
 * 0 as the reference leaves them.  The solve is the same: on the device a
 * support is already a masked row and column of the node-block matrix. */
int  ufem_sys_set_fbc(ufem_sys *s, int fbc);
/* the reference's nDOFfld (fem_dofs, fem/fem_sol.c:1263-1391): 7 entries per
 * node, the 1-based K positions of ux, uy, uz (node-major), 0 elsewhere and,
 * with -fbc, at zero-valued supports; returns the active DOFs (nDOFAct) */
long ufem_sys_dof_map(const ufem_sys *s, long *map);
int  ufem_sys_get_u(ufem_sys *s, double *u);
int  ufem_sys_set_u(ufem_sys *s, const double *u);
int  ufem_sys_get_F(ufem_sys *s, double *F);
/* K as sorted scalar CSR (0-based), explicit zeros included */
int  ufem_sys_get_K(ufem_sys *s, long *indptr, int *cols, double *vals);
int  ufem_sys_spmv(ufem_sys *s, const double *x, double *y);   /* host in, host out */
/* y = K x from the FP32 copy of K that the mixed-precision solvers use
 * (no counterpart in the reference): the copy rounds every value to FP32
 * (with rem, plus its remainder rounded to bf16: 48-bit values) and, for 3x3
 * node blocks, adds to each diagonal block the symmetric part of its block
 * row's summed rounding errors, so rigid translations, which a stiffness
 * matrix's rows cancel, stay cancelled; FP64 vectors and sums (the
 * FP32-vector solver, UFEM_SOLVER_MIXED, uses the plain rounded values).
 * UFEM_NO_CORR32 set in the environment drops the correction. */
int  ufem_sys_spmv_f32(ufem_sys *s, const double *x, double *y, int rem);
/* average device time (HIP events on the system's stream) of `reps`
 * launches of one kernel, and the bytes one launch moves by the algorithmic
 * count.  mode 0: y = K x (FP64); 1: y = K x (FP32 copy of K); 2: the PCG
 * phase-A kernel (p' = z + beta p, q = K p', (p', q)); 3: q = K p with (p, q);
 * 4: w = K u with (w, u), which dominates the default solve; 5: a device copy
 * of one vector; 6: a plain streaming read of K's stored values and columns
 * (the read rate the SpMV's matrix stream can reach; bytes as stored);
 * 7: the same read with non-temporal loads. */
int  ufem_sys_time_spmv(ufem_sys *s, int reps, int mode, double *ms, double *bytes);
/* phase device times of the last calls: fill_K, add_disps, solve (ms) */
void ufem_sys_timings(const ufem_sys *s, double t[3]);

/* ---- uFEM's input file ------------------------------------------------------
 * femReadInput (fem/fem_io.c:380-640) for the linear-static brick path:
 * solution control, real sets, material sets (type 1), nodes, elements
 * (type 9), gravity (none), nodal loads (types 1, 3), element loads (none).
 * preparsed = 1: element node lists, material / real sets and load nodes are
 * 0-based positions (the fem binary's -p); preparsed = 0: they are IDs,
 * renumbered to positions through one hash table per ID list, where
 * femRenumIndex (fem/fem_io.c:50-80) searches each list linearly per entry;
 * the first item carrying an ID wins, as there, and an ID no item carries is
 * AF_ERR_VAL.  *rv receives the AF_* code; NULL on error. */
typedef struct ufem_model ufem_model;
ufem_model *ufem_model_read(const char *fname, int preparsed, int *rv);
void ufem_model_free(ufem_model *m);
/* the model as a mesh for ufem_sys_create (pointers into the model, valid
 * until ufem_model_free), E and nu from each element's material set */
int  ufem_model_mesh(const ufem_model *m, ufem_mesh *mesh);
/* IDs as read (node, element) and the renumbered per-element material and
 * real-set positions */
const long *ufem_model_node_ids(const ufem_model *m);
const long *ufem_model_elem_ids(const ufem_model *m);
const long *ufem_model_elem_mat(const ufem_model *m);
const long *ufem_model_elem_rset(const ufem_model *m);
long ufem_model_sol_id(const ufem_model *m);
long ufem_model_sol_dep(const ufem_model *m);

/* ---- two-level correction ---------------------------------------------------
 * No counterpart in the reference (its solvers precondition with the diagonal
 * or SSOR).  For a box mesh of W x H x D nodes (node x + W (y + H z), element
 * ex + (W-1) (ey + (H-1) ez) with the brick's node order, regular spacing),
 * a coarse node grid every cx, cy, cz elements; P interpolates trilinearly
 * from it, and z_c = P A^-1 P^T r with A = P^T K P (coarse DOFs whose
 * coincident fine DOF is fixed are left out).  A is assembled on the device
 * from the current moduli and inverted (dense) at the start of each
 * UFEM_SOLVER_TWOLEVEL solve, and by the learned solvers when opts.coarse is
 * set.  The factors must divide W-1, H-1, D-1; at most 4096 bricks per cell. */
int  ufem_sys_set_coarse(ufem_sys *s, int W, int H, int D, int cx, int cy, int cz);
long ufem_sys_coarse_size(const ufem_sys *s);     /* coarse DOFs, 0 without a coarse grid */
/* the coarse matrix (inverse = 0) or its inverse (1) for the current moduli,
 * dense row-major [size][size], identity rows at the left-out DOFs */
int  ufem_sys_coarse_matrix(ufem_sys *s, int inverse, double *A);
/* inverse of a dense SPD n x n matrix (row-major) on the device, by block
 * Gauss-Jordan without pivoting -- the coarse solver's factorisation */
int  ufem_gpu_dense_spd_inverse(long n, const double *A, double *Ainv);

/* ---- many samples in one persistent launch ------------------------------
 * The reference solves Monte-Carlo samples one at a time through its plugin
 * (monte_solution -> monte_linear_solver, fem/fem_mont.c:480-511, :694-855).
 * A batch takes n meshes; those that share connectivity, coordinates, loads
 * and supports share one device pattern, every sample keeps its own E, nu,
 * K values and Krylov vectors.  One launch solves all samples with Jacobi
 * PCG (x0 = 0, ||r|| <= eps ||f|| each), a workgroup per sample at a time,
 * samples dealt largest first. */
typedef struct ufem_batch ufem_batch;
ufem_batch *ufem_batch_create(int n, const ufem_mesh *meshes);
void ufem_batch_destroy(ufem_batch *b);
int  ufem_batch_nclasses(const ufem_batch *b);
long ufem_batch_ndof(const ufem_batch *b, int i);
/* element matrices, loads and supports of every sample; device ms out */
int  ufem_batch_assemble(ufem_batch *b, int npts, double *ms);
/* pick the launch shape (workgroup size, workgroups, team size of the
 * largest meshes) by timing each candidate on a budget of iterations per
 * sample; cached in `cache` (a text file; NULL: no cache) */
int  ufem_batch_autotune(ufem_batch *b, const char *cache, int budget_iters, int *wg, int *grid);
void ufem_batch_shape(const ufem_batch *b, int *wg, int *grid, int *team_big);
void ufem_batch_set_shape(ufem_batch *b, int wg, int grid, int team_big);
/* workgroups that solved each sample in the last solve */
int  ufem_batch_teams(const ufem_batch *b, int *team);
/* every team wait of the launch gives up after wait_ms (<= 0: 10 s), and
 * the host stops a solve that has run timeout_ms (<= 0: never) -- the
 * kernel then ends normally, with status 4 for the samples it stopped */
void ufem_batch_set_limits(ufem_batch *b, double wait_ms, double timeout_ms);
/* stop a running ufem_batch_solve (call from another host thread) */
void ufem_batch_abort(ufem_batch *b);
/* no counterpart in the reference: mode 1 solves every sample with FP64
 * vectors over an FP32 copy of its K (each diagonal 3x3 block carries its
 * block row's summed rounding errors, so rigid translations stay exact),
 * mode 2 with the bf16 remainder K - fl32(K) beside each value (48-bit
 * values), mode 0 (the default) with K in FP64;
 * the residual is recomputed in FP64 (r = f - K x, the direction kept) when it
 * has fallen 100-fold, every 500 iterations and before a sample is declared
 * converged.  The copies are built at the next autotune or solve. */
int  ufem_batch_set_mixed(ufem_batch *b, int mode);
/* solve all samples; per-sample outputs may be NULL; maxit <= 0: largest n.
 * sample_ms: device wall-clock time from a sample's start to its end.
 * status per sample: 1 converged, 2 breakdown, 3 maxit, 4 stopped (a team
 * wait gave up or the host stopped the solve), 0 never started; AF_ERR
 * unless every sample converged */
int  ufem_batch_solve(ufem_batch *b, double eps, long maxit, double *total_ms, long *iters,
                      double *relres, int *status, double *sample_ms);
int  ufem_batch_get_u(ufem_batch *b, int i, double *u);

/* ---- learned preconditioner -----------------------------------------------
 * No counterpart in the reference: an accelerator for the Jacobi-PCG solve
 * of a box mesh of W x H x D nodes numbered x fastest (node = x + W (y + H z),
 * DOF 3 node + dir - 1, as ufem_sys numbers them).  Flexible CG with
 *   z = dinv r + mask dsq N(mask dsq r),  dsq = sqrt(dinv),
 * where N is a symmetric L-level 3-D convolution U-Net (3x3x3 taps, no
 * bias; L = 2..8, 3 by default): C0 3->16 channels, C1 16->32, C2 32->64 and
 * C3.. 64->64 with stride 2 (levels of (n + 1) / 2 nodes per axis),
 * N(v) = C0^T (g0 h0 + C1^T (g1 h1 + ... + C_{L-1}^T (g_{L-1} h_{L-1}))),
 * h0 = C0 v, h_l = C_l h_{l-1}, and per-node, per-channel gates
 * g_l = exp(b_l - (1 + a_l) ebar_l) of the log-modulus log(E / 30 GPa) (node
 * average of the bricks, then 3x3x3 means per coarser level).  N is linear,
 * symmetric and positive semi-definite, so M is SPD.  Jacobi PCG takes over
 * if a direction fails to descend.  Parameters (f32, packed): W0 [27][16][3],
 * W1 [27][32][16], W2 [27][64][32], W3.. [27][64][64], then a_l, b_l per
 * level (its channel count each); taps t = 9 (dz+1) + 3 (dy+1) + (dx+1).
 * The coarse-first form (fine = 0) has no node-grid level: C0 is a stride-2
 * 3 -> 32 convolution from the node grid, C1 32 -> 64, C2.. 64 -> 64, level
 * l on the node grid coarsened l + 1 times (ebar pooled once more), and
 * N(v) = C0^T (g0 h0 + ...) back on the node grid; it skips the node-grid
 * activations, the largest of the network.  Weights enter the matrix cores
 * as bf16, activations as f32 (exact bf16 splits). */
typedef struct ufem_net ufem_net;
ufem_net *ufem_net_create(int W, int H, int D, int batch);  /* 3 levels; batch: training samples per step */
ufem_net *ufem_net_create_levels(int W, int H, int D, int batch, int levels);   /* fine = 1 */
ufem_net *ufem_net_create_form(int W, int H, int D, int batch, int levels, int fine);
/* width: at most this many channels per layer (a multiple of 4, 4..64; the
 * forms above use 64) */
ufem_net *ufem_net_create_width(int W, int H, int D, int batch, int levels, int fine, int width);
void ufem_net_destroy(ufem_net *N);
long ufem_net_nparams(void);                 /* of the 3-level network: 70,640 */
long ufem_net_nparams_levels(int levels);    /* -1 outside 2..8 */
long ufem_net_nparams_form(int levels, int fine);
long ufem_net_nparams_width(int levels, int fine, int width);   /* -1 out of range */
int  ufem_net_fine(const ufem_net *N);
int  ufem_net_levels(const ufem_net *N);
int  ufem_net_init(ufem_net *N, unsigned long seed);   /* W ~ N(0, 1/(27 cin)), a = 0, b = -3 */
int  ufem_net_set_weights(ufem_net *N, const float *w);
int  ufem_net_get_weights(ufem_net *N, float *w);
/* y = N(x) for nb samples: x [nb][V][3], ebar0 [nb][V] (node log-moduli),
 * y [nb][V][3] (V = W H D) */
int  ufem_net_forward(ufem_net *N, int nb, const float *x, const float *eb0, float *y);
/* gradient of sum <dy, N(x)> with respect to the packed parameters */
int  ufem_net_backward(ufem_net *N, int nb, const float *x, const float *eb0, const float *dy, float *grad);
/* attach to a system and its current E per element (mask of fixed DOFs,
 * gate inputs); required by collect, train and UFEM_SOLVER_LEARNED */
int  ufem_net_bind(ufem_net *N, ufem_sys *s, const double *E);
/* room for `cap` training pairs on the device */
int  ufem_net_reserve(ufem_net *N, long cap);
long ufem_net_npairs(const ufem_net *N);
/* solve the bound system with Jacobi PCG to eps (relative), snapshot (x, r)
 * every `every` iterations, keep the field (FP32 K, dinv, gate inputs) and
 * add up to `pairs` evenly spread pairs (r, e = u - x); returns the pairs
 * added (< 0: AF_* error) */
long ufem_net_collect(ufem_net *N, int pairs, double eps, int every, ufem_solve_info *info);
/* pair i as stored (scaled by 1 / rms(dsq r)): network input v [V][4],
 * r and e [3 V], e.r, and the index of its material field */
int  ufem_net_get_pair(ufem_net *N, long i, float *v, float *rr, float *ee, double *er, int *field);
/* energy loss (e - z)^T (r - K z) / (e^T r) of pairs [i0, i0 + cnt) */
int  ufem_net_eval(ufem_net *N, long i0, long cnt, double *loss);
/* Adam (0.9, 0.999) on the mean energy loss of `batch` random pairs per
 * step, learning rate decaying from lr to lr / 10 (cosine); loss[steps] and
 * device ms may be NULL */
int  ufem_net_train(ufem_net *N, long steps, int batch, double lr, unsigned long seed, double *loss,
                    double *ms);
/* one convolution kernel on host data, for tests: y = conv(x) with
 * w [27][cout][cin] (mode 0 stride 1, 1 stride 2, 2 stride-2 transpose; flip:
 * taps reversed; x_f32: 1 f32 activations, 2 f32 activations stored in
 * the split format, 0 bf16), dw = its weight gradient for dy (modes 0, 1;
 * x_f32 0 or 1) */
int  ufem_net_probe_conv(int mode, int flip, int x_f32, int nb, const int *gi, const int *go, int cin,
                         int cout, const float *x, const float *w, float *y, const float *dy, float *dw);
/* y = N(v) for the bound field with the solver's FP64 inference (FP64
 * activations and sums, bf16 weights widened exactly, f32 gates): v, y
 * [D][H][W][3] on the node grid, host FP64; the coarse-first form with 16 or
 * 32 channels per level (no counterpart in the reference: the network
 * preconditions its femEqsCGwJ, fem/fem_eqs.c:88-237) */
int  ufem_net_apply64(ufem_net *N, const double *v, double *y);
/* average device time (us) of each of the 2 L inference stages on the bound
 * system's grid: C0, C1, .., C_{L-1} (with the gate product), C_{L-1}^T, ..,
 * C1^T, C0^T */
int  ufem_net_time_stages(ufem_net *N, int reps, double *us);
/* the same for the solver's FP64 inference (coarse-first form, 16 or 32
 * channels): C0 .. C_{L-1}, C_{L-1}^T .. C1^T, and C0^T on its own (in a
 * solve it runs inside the flexible-CG combine) */
int  ufem_net_time_stages64(ufem_net *N, int reps, double *us);
/* average device time (us) of `reps` launches of one such convolution */
int  ufem_net_time_conv(int mode, int flip, int x_f32, int nb, const int *gi, const int *go, int cin, int cout,
                        int reps, double *us);

/* element stiffness matrices of n bricks (xyz: n*24 in node order), dense
 * row-major 24x24 each -- e009_stiff Mode=AF_NO, fem/fem_e009.c:261-635 */
int  ufem_gpu_element_stiffness(long n, const double *xyz, const double *E, const double *nu,
                                int npts, double *Ke);

/* ---- the drop-in's device copies of row-slot matrices (fem_dropin.c) ------
 * policy 0 (default): every drop-in call reads the host matrix once --
 *   fingerprints its columns (a changed pattern rebuilds the copy) and
 *   uploads its values -- so any host change between calls is seen;
 * policy 1: a copy stays valid between calls until ufem_gpu_mark_dirty
 *   (values changed on the host: femMatPutAdd, femMatSetZeroRow/Col, the
 *   supports of fem_add_disps) or ufem_gpu_release (matrix freed, pattern
 *   changed) is called for the matrix. */
void ufem_gpu_set_mirror_policy(int policy);
void ufem_gpu_mark_dirty(const tMatrix *a);
void ufem_gpu_release(const tMatrix *a);
/* shape of a matrix's device copy and its host costs: info[0] block size
 * (3: node blocks, 1: scalar), [1] rows, [2] filled slots, [3] ms of the
 * last build, [4] ms of the last refresh (0 when a valid copy was used) */
int  ufem_gpu_mirror_info(const tMatrix *a, double *info);

/* fem_fill_K(AF_NO) (fem/fem_sol.c:897-989) for a model of bricks (element
 * type 9) with isotropic material (type 1): K += sum of the element
 * matrices, computed on the GPU (npts 2 or 3 Gauss points per axis), added
 * into the row-slot K that fem_sol_alloc made (frompos/defpos from femKhit)
 * as femMatPutAdd(FEM_ADD) adds: existing columns take the sum, new ones
 * fill the row's first empty slots (AF_ERR_BIG when none is left).  m gives
 * node coordinates, element nodes (0-based positions, the brick's order), E
 * and nu per element; its load list is not used.  dof_map is the
 * reference's nDOFfld: known_dofs entries per node, entries 0..2 the 1-based
 * K positions of ux, uy, uz, 0 for a DOF not in K (-fbc, unused node).
 * The device system is kept for the next call on the same connectivity. */
int  ufem_gpu_fill_K(tMatrix *K, const ufem_mesh *m, const long *dof_map, long known_dofs, int npts);
/* of the last ufem_gpu_fill_K: host set-up, device assembly, write-back (ms) */
void ufem_gpu_fill_K_timings(double t[3]);
/* frees the device system ufem_gpu_fill_K keeps */
void ufem_gpu_fill_K_release(void);

/* stopping rule used by the drop-in femEqsCGwJ / femEqsBiCCSwJ (default 0) */
void ufem_gpu_set_stop_rule(int rule);
/* details of the last drop-in solve */
void ufem_gpu_last_solve(ufem_solve_info *info);
/* text of the last error */
const char *ufem_gpu_last_error(void);
/* device name, CU count, HBM bytes */
int  ufem_gpu_device_info(char *name, int name_len, int *n_cu, long *hbm_bytes);

#ifdef __cplusplus
}
#endif
#endif
