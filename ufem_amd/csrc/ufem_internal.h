/*
 * ufem_internal.h -- host-side state shared by the C sources of
 * libufem_gpu.so (not installed; the public API is include/ufem_gpu.h and
 * include/ufem_fem_math.h).
 */
#ifndef UFEM_INTERNAL_H
#define UFEM_INTERNAL_H

#include <stddef.h>

#include "ufem_gpu.h"
#include "ugk.h"

#define UFEM_KNOWN_DOFS 3

/* colour-permuted copy of a matrix for CG with the SSOR preconditioner
 * (ufem_ssor.c): built for one matrix and pattern generation, refilled with
 * its values at every solve */
typedef struct ufem_ssor {
  const ugk_mat *src;
  unsigned long gen;
  long src_slots;
  ugk_mat Kc;
  int *d_perm;              /* new block row -> old block row */
  int *d_dpos;              /* slot of each row's diagonal block in Kc */
  int ncol;
  long *cb;                 /* host, ncol + 1: first row of each colour */
} ufem_ssor;
int ufem_ssor_prepare(const ugk_mat *K, ufem_ssor **S, void *stream);
void ufem_ssor_free(ufem_ssor *S);
/* two-level correction (ufem_coarse.c): A = P^T K P assembled and inverted
 * for the current element moduli */
int ufem_coarse_build(ugk_coarse *C, void *stream);
void ufem_coarse_free(ugk_coarse *C);

/* device vectors of one Krylov workspace (FP64 plus the FP32 inner set) */
typedef struct {
  long n;
  double *r, *z, *p[2], *q, *dinv;
  double *rhat, *phat, *v, *s, *shat, *t;          /* BiCGStab            */
  float *r32, *z32, *p32[2], *q32, *x32, *dinv32;  /* mixed precision     */
  double *fP, *fQ;                                 /* flexible CG rings   */
  int fm;                                          /* ring slots          */
  ufem_ssor *ssor;                                 /* SSOR colour order   */
  ugk_coarse *cc;                                  /* two-level correction */
  double *part;                                    /* reduction partials  */
  ugk_state *st;                                   /* device scalars      */
  ugk_state *st_host;                              /* pinned mirror       */
  /* optional: called by the chunk loop with the iterations enqueued so far
   * (0 before the first chunk), on the solver's stream */
  int (*hook)(void *ctx, long it, void *stream);
  void *hook_ctx;
} ufem_work;

struct ufem_sys {
  long n_nodes, n_elems, n;
  ugk_mat K;
  void *stream;
  void *ev[6];
  /* mesh on the device */
  double *d_x, *d_y, *d_z, *d_E, *d_nu;
  int *d_enodes;
  int n_colors;
  long *color_off;          /* host, n_colors + 1 */
  int *d_elist;             /* elements grouped by colour */
  int maxb;                 /* longest block row */
  /* nodal loads / supports, input order */
  long n_nl;
  long *nl_node, *nl_type, *nl_dir;
  double *nl_val;
  double *F_host;           /* F as the solve takes it (after add_disps)     */
  double *F_load;           /* nodal forces alone, as add_loads left them    */
  long n_fixed;             /* distinct fixed DOFs, first-occurrence order */
  int *fixed_host, *d_fixed;
  double *fix_val;          /* prescribed value of each fixed DOF          */
  int any_prescribed;       /* some fixed DOF has a non-zero value         */
  long n_react;             /* reaction slots, one per support entry */
  long *react_node, *react_dof, *react_fix;
  double *react_init;       /* -F at the DOF when the support was applied */
  double *d_saved_rows;
  int *d_saved_cols;
  double *d_Rfix;
  int filled, loaded, bc_applied;
  int fbc;                  /* result file as with -fbc: zero supports left out   */
  long *node_id, *elem_id;  /* IDs for the result file (NULL: 1-based positions) */
  /* vectors */
  double *d_F, *d_u;
  ufem_work w;
  double t_fill_ms, t_bc_ms, t_solve_ms;
};

/* workspace helpers (ufem_solve.c) */
int  ufem_work_alloc(ufem_work *w, long n, int with_bicgs, int with_mixed);
void ufem_work_free(ufem_work *w);
/* solvers on a device matrix; b, x device vectors of length n (ufem_solve.c) */
int ufem_solve_device(const ugk_mat *K, long n, const double *d_b, double *d_x, ufem_work *w,
                      void *stream, void *ev0, void *ev1, const ufem_solve_opts *o,
                      ufem_solve_info *info);

/* Jacobi PCG with its workspace in the fixed device array (allocates nothing) */
int ufem_solve_device_noalloc(const ugk_mat *K, long n, const double *d_b, double *d_x, void *stream,
                              const ufem_solve_opts *o, ufem_solve_info *info);

/* learned preconditioner inside flexible CG (ufem_net.c) */
long ufem_net_bound_n(const ufem_net *N);           /* DOFs of the bound system, 0 if none */
unsigned short *ufem_net_input(ufem_net *N);       /* split format [V][3][4], channel 3 stays 0 */
const float *ufem_net_mask(const ufem_net *N);     /* 0 at fixed DOFs, else 1             */
const float *ufem_net_output(const ufem_net *N);   /* f32 [V][4]                          */
int ufem_net_prepare(ufem_net *N, void *stream);   /* gates of the bound field            */
int ufem_net_infer(ufem_net *N, const int *d_skip, void *stream);
int ufem_net_infer_head(ufem_net *N, const int *d_skip, void *stream);  /* all stages but the last */
int ufem_net_last_stage(const ufem_net *N, ugk_conv_desc *d);          /* 1: *d is y = C0^T t0 (T2) */
int ufem_net_f64_ok(const ufem_net *N);             /* the form has an FP64 inference path */
double *ufem_net_input64(ufem_net *N);              /* FP64 input mask sqrt(dinv) r, [V][3] */
int ufem_net_prepare64(ufem_net *N, void *stream);  /* buffers and FP64 weight copies */
int ufem_net_infer64_head(ufem_net *N, const int *d_skip, void *stream);
int ufem_net_last64(const ufem_net *N, ugk_conv_desc *d);  /* 1: *d is y = C0^T t0 from FP64 t0 */

/* FP32 copy of K (ufem_sys.c): allocates d_vals32 (and, for 3x3 blocks,
 * d_corr32; with rem, d_rem16) on first use and converts the current FP64
 * values */
int ufem_mat_f32(ugk_mat *K, int rem, void *stream);

/* error text (ufem_sys.c) */
void ufem_set_error(const char *fmt, ...);

/* uFEM's preparsed input file as femReadInput holds it (fem/fem_io.c:380-640),
 * every index 0-based (ufem_input.c).  Sets of values (real sets, materials)
 * keep the reference's flat list + start offsets. */
typedef struct {
  long sol_id, sol_dep_on, sol_lin, sol_nonlin, sol_nonlin_steps;
  long rs_len, rs_len_l, *rs_id, *rs_type, *rs_from;
  double *rs_val;
  long mp_len, mp_len_l, *mp_id, *mp_type, *mp_from;
  double *mp_val;
  long n_len, *n_id;
  double *n_x, *n_y, *n_z;
  long e_len, e_len_l, *e_id, *e_type, *e_rs, *e_mp, *e_from, *e_nodes;
  long gr_dir;
  double gr_val;
  long nl_len, *nl_node, *nl_type, *nl_dir;
  double *nl_val;
  long el_len;
  /* the Monte section (fem_monte_read_data, fem/fem_mont.c:866-909):
     inputs first, then outputs */
  long mc_ilen, mc_olen, mc_fail_id;
  long *mc_pos, *mc_type, *mc_item, *mc_sub, *mc_iter;
} ufem_input;
int  ufem_input_read(const char *fname, ufem_input *in);          /* preparsed (-p) */
int  ufem_input_read_form(const char *fname, int preparsed, ufem_input *in);
void ufem_input_free(ufem_input *in);
/* value `type` (repetition `rep`, 0: the non-repeating part) of material set
 * `mp` of type 1: its index in mp_val, -1 when absent (femGetRepValIndex) */
long ufem_input_mat_index(const ufem_input *in, long mp, long type, long rep);
/* E and nu of every element from its material set (femGetMPValPos) */
void ufem_input_elastic(const ufem_input *in, double *E, double *nu);

#endif
