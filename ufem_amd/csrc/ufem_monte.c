/*
 * ufem_monte.c -- uFEM's plugin interface for the Monte probabilistic
 * program (fem/fem_mont.c, built into the reference's libfem.so with
 * USE_MONTE, fem/Makefile:192-193), over the device-resident solve.
 *
 * Monte loads the library, calls monte_init_lib_stuff with the name of a
 * preparsed input file that ends in a section of random variables, asks for
 * the numbers and names of the input and output variables, and then calls
 * monte_solution once per sample with the sampled inputs.  Each sample here
 * updates the device system in place (coordinates, moduli, load and support
 * values; the pattern, colouring and buffers stay from the first sample),
 * refills K, applies loads and supports, and solves with CG and the SSOR
 * preconditioner under the reference's own call (femEqsCGwSSOR with eps =
 * FEM_ZERO / 10000 and 3 n iterations at most, fem/fem_mont.c:480-511).
 * The rows go in greedy colour order on the device rather than the natural
 * order, so the iterates differ from the reference's and the converged
 * solution agrees to the tolerance of the solve.
 *
 * Variable codes (fem/fem_mont.h:30-50) and every quirk of the reference's
 * bookkeeping are kept, so that a Monte set-up reads the same numbers:
 * output k writes ofld[pos[k]] with pos[] counted from the first INPUT row
 * (fem/fem_mont.c:521, :700), displacement sums add DOF-map entries, maxima
 * and minima start from 0.  State is global, as in the reference: one model
 * per process.
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ufem_fem_math.h"
#include "ufem_internal.h"
#include "ufem_monte.h"

enum {
  VT_RS = 0, VT_MAT = 1, VT_N = 2, VT_NLD = 3, VT_EL = 4, VT_NLPOS = 5, VT_DAMP = 6, VT_MAX_D = 7,
  VT_RES_D = 8, VT_RES_R = 9, VT_RES_E = 10, VT_RES_SUM_E = 11, VT_RES_MAX_E = 12, VT_RES_MIN_E = 13,
  VT_RES_FAIL_E = 14, VT_RES_D_MAX = 15, VT_RES_D_MIN = 16, VT_RES_D_SUM = 17, VT_RES_A_MAX = 18,
  VT_RES_FAIL = 19
};

#define KNOWN_DOFS 7          /* DOF-map entries per node (fem/fem_elem.h:43) */
#define RES_RP 20             /* repeating results per point of element 9   */
#define RES_PTS 27            /* results-pass points of element 9           */
#define RES_LEN (14 * 27)     /* result values per element (e009_rvals)     */
#define FEM_ZERO 1e-17        /* fem/fem.h:59 */

/* result types of element 9 in their repeating order (fem/fem_e009.c:879-884,
 * ids fem/fem_eres.h): sigma x y z yz zx xy, epsilon x y z yz zx xy, plastic
 * strains x y z yz zx xy, CR1, PSI */
static const long e9_res_rp[RES_RP] = {7, 8, 9, 11, 12, 10, 13, 14, 15, 17, 18, 16,
                                       73, 74, 75, 77, 78, 76, 37, 38};

static struct {
  int ready;
  ufem_input in;
  double *E, *nu;
  ufem_sys *sys;
  double *u, *R, *res;        /* host copies of the last solution          */
  long *rnode, *rdof;
  long nres;                  /* reaction list: zero-valued supports       */
  long *res_entry;            /* ... their support entries                 */
  double *racc;               /* ... and their values summed over samples  */
  double max_disp, damp_alpha, damp_beta;
  long solves;
  char name[257];
} M;

long monte_dlib_interface_type(void) { return 2; }

void monte_nums_of_vars(char *param, long *ilen, long *olen, long *ffunc)
{
  (void)param;
  *ilen = M.in.mc_ilen;
  *olen = M.in.mc_olen;
  *ffunc = M.ready ? M.in.mc_fail_id : -1;
}

static const char *dir_xyz(long d)
{
  return d == 1 ? "X" : d == 2 ? "Y" : d == 3 ? "Z" : "_";
}
static const char *dir_res(long d, int reac)
{
  static const char *u[6] = {"UX", "UY", "UZ", "RX", "RY", "RZ"};
  static const char *f[6] = {"FX", "FY", "FZ", "MX", "MY", "MZ"};
  return (d >= 1 && d <= 6) ? (reac ? f[d - 1] : u[d - 1]) : "__";
}

/* names as fem/fem_mont.c:158-250 prints them; an item outside the model
 * (undefined in the reference) names "empty" */
char *monte_ivar_name(char *param, long pos)
{
  const ufem_input *in = &M.in;
  long it, sb, ir;
  (void)param;
  if (pos < 0 || pos >= in->mc_ilen) return "empty";
  it = in->mc_item[pos]; sb = in->mc_sub[pos]; ir = in->mc_iter[pos];
  switch (in->mc_type[pos]) {
    case VT_RS:
      if (it < 0 || it >= in->rs_len) return "empty";
      snprintf(M.name, sizeof(M.name), "RS_%li_P%li_I%li", in->rs_id[it], sb, ir);
      return M.name;
    case VT_MAT:
      if (it < 0 || it >= in->mp_len) return "empty";
      snprintf(M.name, sizeof(M.name), "MAT_%li_P%li_I%li", in->mp_id[it], sb, ir);
      return M.name;
    case VT_N:
      if (it < 0 || it >= in->n_len) return "empty";
      snprintf(M.name, sizeof(M.name), "N_%li_%s", in->n_id[it], dir_xyz(sb));
      return M.name;
    case VT_NLD:   /* the name reads the item as a load position */
      if (it < 0 || it >= in->nl_len) return "empty";
      snprintf(M.name, sizeof(M.name), "NLD_N%li_T%li_D%li", in->n_id[in->nl_node[it]], sb, ir);
      return M.name;
    case VT_NLPOS:
      snprintf(M.name, sizeof(M.name), "NLP_%li_POS%li_T%li_D%li", pos + 1, it + 1, sb, ir);
      return M.name;
    case VT_DAMP:
      return it == 1 ? "DAMP_ALPHA" : "DAMP_BETA";
    case VT_MAX_D:
      return "MAX_DISP";
    default:       /* element loads do not exist on this path */
      return "empty";
  }
}

/* fem/fem_mont.c:256-398 */
char *monte_ovar_name(char *param, long pos)
{
  const ufem_input *in = &M.in;
  long it, sb, i;
  (void)param;
  if (pos < 0 || pos >= in->mc_olen) return "empty";
  i = pos + in->mc_ilen;
  it = in->mc_item[i]; sb = in->mc_sub[i];
  switch (in->mc_type[i]) {
    case VT_RES_D: case VT_RES_D_MAX: case VT_RES_D_MIN: case VT_RES_D_SUM: case VT_RES_A_MAX: case VT_RES_R: {
      const char *pre = in->mc_type[i] == VT_RES_D ? "DISP" : in->mc_type[i] == VT_RES_D_MAX ? "DMAX"
                      : in->mc_type[i] == VT_RES_D_MIN ? "DMIN" : in->mc_type[i] == VT_RES_D_SUM ? "DSUM"
                      : in->mc_type[i] == VT_RES_A_MAX ? "ACCMAX" : "REAC";
      if (it < 0 || it >= in->n_len) return "empty";
      snprintf(M.name, sizeof(M.name), "%s_%s_%li", pre, dir_res(sb, in->mc_type[i] == VT_RES_R), in->n_id[it]);
      return M.name;
    }
    case VT_RES_E:
      if (it < 0 || it >= in->e_len) return "empty";
      snprintf(M.name, sizeof(M.name), "ERES_E%li_T%li_I%li", in->e_id[it], sb, in->mc_iter[i]);
      return M.name;
    case VT_RES_SUM_E: snprintf(M.name, sizeof(M.name), "ERES_SUM_T%li", sb); return M.name;
    case VT_RES_MAX_E: snprintf(M.name, sizeof(M.name), "ERES_MAX_T%li", sb); return M.name;
    case VT_RES_MIN_E: snprintf(M.name, sizeof(M.name), "ERES_MIN_T%li", sb); return M.name;
    case VT_RES_FAIL_E: snprintf(M.name, sizeof(M.name), "ERES_FAIL_T%li", sb); return M.name;
    case VT_RES_FAIL: return "RES_FAIL";
    default: return "empty";
  }
}

int monte_clean_lib_stuff(char *param)
{
  (void)param;
  if (M.sys) ufem_sys_destroy(M.sys);
  ufem_input_free(&M.in);
  free(M.E); free(M.nu); free(M.u); free(M.R); free(M.res); free(M.rnode); free(M.rdof); free(M.res_entry);
  free(M.racc);
  memset(&M, 0, sizeof(M));
  M.in.mc_fail_id = -1;
  return 0;
}

/* femReadInput + fem_dofs + fem_sol_alloc + fem_sol_res_alloc
 * (fem/fem_mont.c:424-469).  The device system is made at the first
 * sample, so that the model and its variables can be read on any host.
 * The reaction list holds the supports whose value is 0 here
 * (fem_sol_res_alloc, fem/fem_sol.c:676-722). */
int monte_init_lib_stuff(char *param)
{
  long i, nr = 0, ne, nn, n;
  if (!param || strlen(param) < 1) return -1;
  monte_clean_lib_stuff(param);
  if (ufem_input_read(param, &M.in) != AF_OK) {
    M.in.mc_fail_id = -1;
    return -1;
  }
  ne = M.in.e_len;
  nn = M.in.n_len;
  n = 3 * nn;
  M.E = (double *)malloc(sizeof(double) * (size_t)ne);
  M.nu = (double *)malloc(sizeof(double) * (size_t)ne);
  M.u = (double *)calloc((size_t)n, sizeof(double));
  M.res = (double *)calloc((size_t)ne * RES_PTS * 12, sizeof(double));
  M.R = (double *)calloc((size_t)M.in.nl_len + 1, sizeof(double));
  M.rnode = (long *)calloc((size_t)M.in.nl_len + 1, sizeof(long));
  M.rdof = (long *)calloc((size_t)M.in.nl_len + 1, sizeof(long));
  M.res_entry = (long *)calloc((size_t)M.in.nl_len + 1, sizeof(long));
  M.racc = (double *)calloc((size_t)M.in.nl_len + 1, sizeof(double));
  if (!M.E || !M.nu || !M.u || !M.res || !M.R || !M.rnode || !M.rdof || !M.res_entry || !M.racc) {
    monte_clean_lib_stuff(param);
    return -1;
  }
  for (i = 0; i < M.in.nl_len; i++) {
    if (M.in.nl_type[i] != 1) continue;
    if (M.in.nl_val[i] == 0.0) M.res_entry[M.nres++] = nr;
    nr++;
  }
  M.ready = 1;
  return 0;
}

/* the device system's view of the current model */
static void model_mesh(ufem_mesh *m, long *enodes)
{
  const ufem_input *in = &M.in;
  long e, k;
  for (e = 0; e < in->e_len; e++)
    for (k = 0; k < 8; k++) enodes[8 * e + k] = in->e_nodes[in->e_from[e] + k];
  ufem_input_elastic(in, M.E, M.nu);
  m->n_nodes = in->n_len;
  m->x = in->n_x; m->y = in->n_y; m->z = in->n_z;
  m->n_elems = in->e_len;
  m->enodes = enodes;
  m->E = M.E; m->nu = M.nu;
  m->n_nl = in->nl_len;
  m->nl_node = in->nl_node; m->nl_type = in->nl_type; m->nl_dir = in->nl_dir; m->nl_val = in->nl_val;
}

/* monte_linear_solver (fem/fem_mont.c:480-511): K, F, u from zero, fem_fill_K,
 * fem_add_loads, fem_add_disps, femEqsCGwSSOR(K, F, u, FEM_ZERO/10000,
 * 3 rows), then the results pass for reactions and element results.  The
 * reference never clears its reaction slots between samples (femApplyDisp
 * subtracts the load, fem/fem_loads.c:205; fem_fill_K_React adds K u,
 * fem/fem_sol.c:866), so sample k reports the sum of the reactions of
 * samples 1..k; the same sum is kept here (over the samples that solved). */
/* the device system set to mesh m: created at the first sample, updated in
 * place after; K filled, loads and supports applied */
static int prepare_system(const ufem_mesh *m)
{
  int rv;
  if (!M.sys) {
    M.sys = ufem_sys_create(m);
    rv = M.sys ? AF_OK : AF_ERR;
  } else {
    rv = ufem_sys_update(M.sys, m);
  }
  if (rv != AF_OK) return rv;
  if ((rv = ufem_sys_fill_K(M.sys, 3)) != AF_OK) return rv;
  if ((rv = ufem_sys_add_loads(M.sys)) != AF_OK) return rv;
  return ufem_sys_add_disps(M.sys);
}

/* after the solve (u on the device): host u, reactions added to their sums,
 * element results */
static int sample_results(int need_r, int need_e)
{
  long r;
  int rv;
  M.solves++;
  if ((rv = ufem_sys_get_u(M.sys, M.u)) != AF_OK) return rv;
  if (need_r) {
    if ((rv = ufem_sys_reactions(M.sys, M.R, M.rnode, M.rdof)) != AF_OK) return rv;
    for (r = 0; r < M.nres; r++) {   /* compact to the reaction list */
      M.R[r] = M.R[M.res_entry[r]];
      M.rnode[r] = M.rnode[M.res_entry[r]];
      M.rdof[r] = M.rdof[M.res_entry[r]];
      M.racc[r] += M.R[r];
    }
  }
  if (need_e && (rv = ufem_sys_element_results(M.sys, 3, M.res)) != AF_OK) return rv;
  return AF_OK;
}

static int linear_solution(int need_r, int need_e)
{
  ufem_mesh m;
  ufem_solve_opts o;
  ufem_solve_info info;
  long *enodes = (long *)malloc(sizeof(long) * (size_t)M.in.e_len * 8);
  int rv;
  if (!enodes) return AF_ERR_MEM;
  model_mesh(&m, enodes);
  rv = prepare_system(&m);
  free(enodes);
  if (rv != AF_OK) return rv;
  memset(&o, 0, sizeof(o));
  o.solver = UFEM_SOLVER_SSOR;
  o.rule = UFEM_RULE_REFERENCE;
  o.eps = FEM_ZERO / 10000.0;
  o.maxit = 3 * ufem_sys_ndof(M.sys);
  if ((rv = ufem_sys_solve(M.sys, &o, &info)) != AF_OK) return rv;
  return sample_results(need_r, need_e);
}

/* nDOFfld entry `pos` of the node-major DOF map: ux, uy, uz numbered from 1,
 * the other four DOFs of a node 0 (fem_dofs, fem/fem_sol.c:1263-1391) */
static long dof_map(long pos)
{
  if (pos < 0 || pos >= KNOWN_DOFS * M.in.n_len) return 0;
  return pos % KNOWN_DOFS < 3 ? 3 * (pos / KNOWN_DOFS) + pos % KNOWN_DOFS + 1 : 0;
}
/* femVecGet(&u, pos): 1-based, 0 outside (fem/fem_math.c:811-819) */
static double u_at(long dof)
{
  return (dof < 1 || dof > 3 * M.in.n_len) ? 0.0 : M.u[dof - 1];
}
/* femGetEResVal (fem/fem_dama.h:153 -> femGetRepVal, fem/fem_accs.c:196-279)
 * for element 9: no non-repeating results; value `type` of repetition rep
 * sits at (position in the repeating list) + 20 (rep - 1) of the element's
 * 378 values, 0 past them */
static double eres(long e, long type, long rep)
{
  long i, i_pos = -1, gp, t;
  if (rep <= 0 || e < 0 || e >= M.in.e_len) return 0.0;
  for (i = 0; i < RES_RP; i++)
    if (e9_res_rp[i] == type) { i_pos = i + RES_RP * (rep - 1); break; }
  if (i_pos < 0 || i_pos >= RES_LEN) return 0.0;
  gp = i_pos / RES_RP;
  t = i_pos % RES_RP;
  return t < 12 ? M.res[(e * RES_PTS + gp) * 12 + t] : 0.0;
}

/* monte_fill_ofld_data (fem/fem_mont.c:514-691) */
static void fill_outputs(double *ofld)
{
  const ufem_input *in = &M.in;
  const long il = in->mc_ilen, nd = KNOWN_DOFS * in->n_len;
  long i, j, k;
  for (i = il; i < il + in->mc_olen; i++) {
    double *o = &ofld[in->mc_pos[i - il]];
    const long it = in->mc_item[i], sb = in->mc_sub[i];
    long pos = it * KNOWN_DOFS + sb - 1;
    double v;
    switch (in->mc_type[i]) {
      case VT_RES_D:
        if (pos >= nd || pos < 0) { *o = 0.0; break; }
        *o = u_at(dof_map(pos));
        break;
      case VT_RES_D_MAX:
        if (pos >= nd || pos < 0) break;
        v = u_at(dof_map(pos));
        if (*o < v) *o = v;
        break;
      case VT_RES_D_MIN:
        if (pos >= nd || pos < 0) break;
        v = u_at(dof_map(pos));
        if (*o > v) *o = v;
        break;
      case VT_RES_D_SUM:   /* adds the map entry at the DOF number, not u */
        if (pos >= nd || pos < 0) break;
        *o += (double)dof_map(dof_map(pos));
        break;
      case VT_RES_A_MAX:   /* accelerations exist only in dynamics */
        break;
      case VT_RES_R:
        for (j = 0; j < M.nres; j++)
          if (M.rnode[j] == it && M.rdof[j] == sb) { *o = M.racc[j]; break; }
        break;
      case VT_RES_E:
        *o = eres(it, sb, in->mc_iter[i]);
        break;
      case VT_RES_SUM_E:
        *o = 0.0;
        for (j = 0; j < in->e_len; j++)
          for (k = 0; k <= RES_RP; k++) *o += eres(j, sb, k);
        break;
      case VT_RES_MAX_E:
        *o = 0.0;
        for (j = 0; j < in->e_len; j++)
          for (k = 0; k <= RES_RP; k++)
            if ((v = eres(j, sb, k)) > *o) *o = v;
        break;
      case VT_RES_MIN_E:
        *o = 0.0;
        for (j = 0; j < in->e_len; j++)
          for (k = 0; k <= RES_RP; k++)
            if ((v = eres(j, sb, k)) < *o) *o = v;
        break;
      case VT_RES_FAIL_E:
        *o = 0.0;
        for (j = 0; j < in->e_len && *o == 0.0; j++)
          for (k = 0; k <= RES_RP; k++)
            if (eres(j, sb, k) > 0.0) { *o = 1.0; break; }
        break;
      case VT_RES_FAIL:
        /* fem_asse_max_disp_simple (fem/fem_asse.c:42-46) and
           fem_asse_fail_cond (:271-380): the latter's von Mises test of the
           first element's point 0 compares 0 with the yield stress, which a
           material of type 1 does not hold (femGetMPValPos gives 0), so the
           test fails and so does the sample */
        if (in->mc_fail_id > -1) {
          if (M.max_disp > 0.0) {
            double mx = 0.0;
            for (j = 0; j < 3 * in->n_len; j++)
              if (fabs(M.u[j]) > mx) mx = fabs(M.u[j]);
            if (mx >= fabs(M.max_disp)) *o = 1.0;
          }
          if (in->e_len > 0) *o = 1.0;
        }
        break;
      default:
        break;
    }
  }
}

/* the sampled values of ifld into the model, in row order
 * (fem/fem_mont.c:709-785) */
static void apply_inputs(const double *ifld)
{
  ufem_input *in = &M.in;
  long i, j;
  for (i = 0; i < in->mc_ilen; i++) {
    const long it = in->mc_item[i], sb = in->mc_sub[i], ir = in->mc_iter[i];
    const double v = ifld[in->mc_pos[i]];
    switch (in->mc_type[i]) {
      case VT_MAT: {   /* femGetRepValIndex, then the value replaced */
        long k = ufem_input_mat_index(in, it, sb, ir);
        if (k > -1 && k < in->mp_len_l) in->mp_val[k] = v;
        break;
      }
      case VT_N:
        if (it < 0 || it >= in->n_len) break;
        if (sb == 1) in->n_x[it] = v;
        else if (sb == 2) in->n_y[it] = v;
        else if (sb == 3) in->n_z[it] = v;
        break;
      case VT_NLD:
        for (j = 0; j < in->nl_len; j++)
          if (in->nl_node[j] == it && in->nl_type[j] == sb && in->nl_dir[j] == ir) { in->nl_val[j] = v; break; }
        break;
      case VT_NLPOS:
        if (it >= 0 && it < in->nl_len && in->nl_type[it] == sb && in->nl_dir[it] == ir) in->nl_val[it] = v;
        break;
      case VT_DAMP:
        if (it == 1) M.damp_alpha = v; else M.damp_beta = v;
        break;
      case VT_MAX_D:
        M.max_disp = fabs(v);
        break;
      default:         /* real sets carry nothing for the brick; no element loads */
        break;
    }
  }
}

static void outputs_needed(int *need_r, int *need_e)
{
  const ufem_input *in = &M.in;
  long i;
  *need_r = *need_e = 0;
  for (i = in->mc_ilen; i < in->mc_ilen + in->mc_olen; i++) {
    const long t = in->mc_type[i];
    if (t == VT_RES_R) *need_r = 1;
    if (t == VT_RES_E || t == VT_RES_SUM_E || t == VT_RES_MAX_E || t == VT_RES_MIN_E || t == VT_RES_FAIL_E)
      *need_e = 1;
  }
}

static void zero_outputs(double *ofld)
{
  const ufem_input *in = &M.in;
  long i;
  for (i = in->mc_ilen; i < in->mc_ilen + in->mc_olen; i++) ofld[in->mc_pos[i - in->mc_ilen]] = 0.0;
}

static int check_if_type(long if_type)
{
  if (if_type == 1 || if_type == 2) {
    ufem_set_error("monte_solution: if_type %ld (%s) is not supported", if_type,
                   if_type == 1 ? "price" : "dynamics");
    return AF_ERR_TYP;
  }
  return AF_OK;
}

/* monte_solution (fem/fem_mont.c:694-855): if_type 0 is the linear static
 * solution; 1 (price of the structure, printed before the reference exits)
 * and 2 (dynamics) return AF_ERR_TYP */
int monte_solution(char *param, double *ifld, double *ofld, long if_type)
{
  int rv, need_r, need_e;
  (void)param;
  if (!M.ready) { ufem_set_error("monte_init_lib_stuff first"); return -1; }
  zero_outputs(ofld);
  apply_inputs(ifld);
  if ((rv = check_if_type(if_type)) != AF_OK) return rv;
  outputs_needed(&need_r, &need_e);
  if ((rv = linear_solution(need_r, need_e)) != AF_OK) return rv;
  fill_outputs(ofld);
  return AF_OK;
}

/* n samples in one batched launch: sample k's inputs at ifld + k ilen, its
 * output field at ofld + k olen.  The samples are applied to the model in
 * order, as n calls of monte_solution would, and solved together by the
 * batch's Jacobi PCG to ||r|| <= 1e-15 ||f|| (3 n iterations at most); the
 * reactions and element results of each sample then follow on the device
 * system in sample order.  Models with a support of non-zero value, or
 * inputs that set one, are solved one sample at a time. */
#define MC_BATCH_EPS 1e-15
int monte_solution_batch(char *param, long n, const double *ifld, double *ofld, long if_type, int *status)
{
  ufem_input *in = &M.in;
  const long il = in->mc_ilen, ol = in->mc_olen, ne = in->e_len, nn = in->n_len, nl = in->nl_len;
  ufem_mesh *meshes = NULL;
  ufem_batch *b = NULL;
  double *snap = NULL;
  long *enodes = NULL, *iters = NULL, k, i, sequential = 0;
  int *st = NULL, rv = AF_OK, need_r, need_e;
  const size_t per = (size_t)(2 * ne + 3 * nn + nl);
  if (!M.ready) { ufem_set_error("monte_init_lib_stuff first"); return -1; }
  if (n < 1) return AF_OK;
  for (i = il; i < il + ol; i++)
    if (in->mc_pos[i - il] < 0 || in->mc_pos[i - il] >= ol) {
      ufem_set_error("batch: output slot %ld outside the field of %ld", in->mc_pos[i - il], ol);
      return AF_ERR_SIZ;
    }
  for (i = 0; i < il; i++)
    if (in->mc_pos[i] < 0 || in->mc_pos[i] >= il) {
      ufem_set_error("batch: input slot %ld outside the field of %ld", in->mc_pos[i], il);
      return AF_ERR_SIZ;
    }
  if ((rv = check_if_type(if_type)) != AF_OK) return rv;
  for (i = 0; i < nl; i++)
    if (in->nl_type[i] == 1 && in->nl_val[i] != 0.0) sequential = 1;
  for (i = 0; i < il; i++)
    if ((in->mc_type[i] == VT_NLD || in->mc_type[i] == VT_NLPOS) && in->mc_sub[i] == 1) sequential = 1;
  if (sequential || n == 1) {
    for (k = 0; k < n; k++) {
      int r = monte_solution(param, (double *)(ifld + k * il), ofld + k * ol, if_type);
      if (status) status[k] = r;
      if (r != AF_OK && rv == AF_OK) rv = r;
    }
    return rv;
  }
  outputs_needed(&need_r, &need_e);
  meshes = (ufem_mesh *)calloc((size_t)n, sizeof(ufem_mesh));
  snap = (double *)malloc(sizeof(double) * per * (size_t)n);
  enodes = (long *)malloc(sizeof(long) * (size_t)ne * 8);
  iters = (long *)calloc((size_t)n, sizeof(long));
  st = (int *)calloc((size_t)n, sizeof(int));
  if (!meshes || !snap || !enodes || !iters || !st) { rv = AF_ERR_MEM; goto out; }
  for (k = 0; k < n; k++) {   /* model of sample k, as monte_solution k would see it */
    ufem_mesh m;
    double *E = snap + per * (size_t)k, *nu = E + ne, *x = nu + ne, *y = x + nn, *z = y + nn, *v = z + nn;
    apply_inputs(ifld + k * il);
    model_mesh(&m, enodes);
    memcpy(E, M.E, sizeof(double) * (size_t)ne);
    memcpy(nu, M.nu, sizeof(double) * (size_t)ne);
    memcpy(x, in->n_x, sizeof(double) * (size_t)nn);
    memcpy(y, in->n_y, sizeof(double) * (size_t)nn);
    memcpy(z, in->n_z, sizeof(double) * (size_t)nn);
    if (nl > 0) memcpy(v, in->nl_val, sizeof(double) * (size_t)nl);
    m.E = E; m.nu = nu; m.x = x; m.y = y; m.z = z; m.nl_val = v;
    meshes[k] = m;
  }
  if (!(b = ufem_batch_create((int)n, meshes))) { rv = AF_ERR; goto out; }
  if ((rv = ufem_batch_assemble(b, 3, NULL)) != AF_OK) goto out;
  /* per-sample status 1 = converged; a launch that failed leaves all 0 */
  (void)ufem_batch_solve(b, MC_BATCH_EPS, 3 * 3 * nn, NULL, iters, NULL, st, NULL);
  rv = AF_OK;
  for (k = 0; k < n; k++) {
    double *o = ofld + k * ol;
    int r = st[k] == 1 ? AF_OK : AF_ERR;
    zero_outputs(o);
    if (r == AF_OK) r = prepare_system(&meshes[k]);
    if (r == AF_OK) r = ufem_batch_get_u(b, (int)k, M.u);
    if (r == AF_OK) r = ufem_sys_set_u(M.sys, M.u);
    if (r == AF_OK) r = sample_results(need_r, need_e);
    if (r == AF_OK) fill_outputs(o);
    if (status) status[k] = r;
    if (r != AF_OK && rv == AF_OK) rv = r;
  }
out:
  if (b) ufem_batch_destroy(b);
  free(meshes); free(snap); free(enodes); free(iters); free(st);
  return rv;
}
