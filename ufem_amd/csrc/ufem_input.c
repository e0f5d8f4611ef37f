/*
 * ufem_input.c -- reader of uFEM's preparsed input file (fem/feminput.5, the
 * form femReadInput takes with femPreparsedData set, fem/fem_io.c:380-640),
 * for the linear-static brick path: solution control, real sets, material
 * sets, nodes, elements, gravity, nodal and element loads
 * (femReadInputLoads, fem/fem_io.c:86-175), the time-dependent array data
 * and dynamic loads (femReadRepeatData :178-298, femReadInputDynLoads
 * :301-360, accepted only when empty) and the Monte section
 * (fem_monte_read_data, fem/fem_mont.c:866-909).
 *
 * Preparsed files carry 0-based positions where the reference would carry
 * IDs (element material and real-set fields, element node lists, load
 * nodes), so nothing is renumbered (femRenumIndex returns at once,
 * fem/fem_io.c:55); the form with IDs (the fem binary without -p) is
 * renumbered by hash lookups (renumber_ids below).  Sizes that disagree with their item counts return
 * AF_ERR_SIZ, fewer than 2 nodes or no element AF_ERR_SML, as the reference
 * does; a short or unreadable file returns AF_ERR_IO.
 */
#define _POSIX_C_SOURCE 200809L
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ufem_fem_math.h"
#include "ufem_internal.h"

#define AF_ERR_SML -10

/* material type 1 (fem/fem_m001.c:35-50): DENS, EX, NU, ALPHA, KXX, C,
 * PRICE, then repeating DENS, EX, NU, ALPHA, KXX, C; type ids from
 * fem/fem_mat.h:40-78 */
#define MAT_EX 2
#define MAT_NU 5
static const long mat1_val[7] = {1, 2, 5, 38, 37, 39, 36};
static const long mat1_val_rp[6] = {1, 2, 5, 38, 37, 39};

static int rd_l(FILE *f, long *v) { return fscanf(f, "%li", v) == 1; }
static int rd_d(FILE *f, double *v) { return fscanf(f, "%lf", v) == 1; }

static long *lalloc(long n) { return (long *)calloc((size_t)(n > 0 ? n : 1), sizeof(long)); }
static double *dalloc(long n) { return (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

void ufem_input_free(ufem_input *in)
{
  if (!in) return;
  free(in->rs_id); free(in->rs_type); free(in->rs_from); free(in->rs_val);
  free(in->mp_id); free(in->mp_type); free(in->mp_from); free(in->mp_val);
  free(in->n_id); free(in->n_x); free(in->n_y); free(in->n_z);
  free(in->e_id); free(in->e_type); free(in->e_rs); free(in->e_mp); free(in->e_from); free(in->e_nodes);
  free(in->nl_node); free(in->nl_type); free(in->nl_dir); free(in->nl_val);
  free(in->mc_pos); free(in->mc_type); free(in->mc_item); free(in->mc_sub); free(in->mc_iter);
  memset(in, 0, sizeof(*in));
}

/* "id type count" lines, then the total and the values: counts turn into
 * start offsets (fem/fem_io.c:411-455, :481-525) */
static int read_sets(FILE *f, long *len, long **id, long **type, long **from, long *len_l, double **val)
{
  long i, sum = 0, c;
  if (!rd_l(f, len)) return AF_ERR_IO;
  if (*len <= 0) { *len = 0; *len_l = 0; return AF_OK; }
  *id = lalloc(*len); *type = lalloc(*len); *from = lalloc(*len);
  if (!*id || !*type || !*from) return AF_ERR_MEM;
  for (i = 0; i < *len; i++) {
    if (!rd_l(f, &(*id)[i]) || !rd_l(f, &(*type)[i]) || !rd_l(f, &c)) return AF_ERR_IO;
    (*from)[i] = sum;
    sum += c;
  }
  if (!rd_l(f, len_l)) return AF_ERR_IO;
  if (*len_l != sum) {
    ufem_set_error("set values: %ld given, %ld counted", *len_l, sum);
    return AF_ERR_SIZ;
  }
  if (!(*val = dalloc(*len_l))) return AF_ERR_MEM;
  for (i = 0; i < *len_l; i++)
    if (!rd_d(f, &(*val)[i])) return AF_ERR_IO;
  return AF_OK;
}

static int read_body(FILE *f, ufem_input *in)
{
  long i, sum, c, trans_num, dyn_num, dyn0;
  double dyn_stp;
  int rv;
  if (!rd_l(f, &in->sol_id) || !rd_l(f, &in->sol_dep_on) || !rd_l(f, &in->sol_lin) ||
      !rd_l(f, &in->sol_nonlin) || !rd_l(f, &in->sol_nonlin_steps))
    return AF_ERR_IO;
  if ((rv = read_sets(f, &in->rs_len, &in->rs_id, &in->rs_type, &in->rs_from, &in->rs_len_l, &in->rs_val)))
    return rv;
  if ((rv = read_sets(f, &in->mp_len, &in->mp_id, &in->mp_type, &in->mp_from, &in->mp_len_l, &in->mp_val)))
    return rv;
  /* nodes (fem/fem_io.c:527-540) */
  if (!rd_l(f, &in->n_len)) return AF_ERR_IO;
  if (in->n_len < 2) { ufem_set_error("fewer than 2 nodes"); return AF_ERR_SML; }
  in->n_id = lalloc(in->n_len);
  in->n_x = dalloc(in->n_len); in->n_y = dalloc(in->n_len); in->n_z = dalloc(in->n_len);
  if (!in->n_id || !in->n_x || !in->n_y || !in->n_z) return AF_ERR_MEM;
  for (i = 0; i < in->n_len; i++)
    if (!rd_l(f, &in->n_id[i]) || !rd_d(f, &in->n_x[i]) || !rd_d(f, &in->n_y[i]) || !rd_d(f, &in->n_z[i]))
      return AF_ERR_IO;
  /* elements (fem/fem_io.c:542-600) */
  if (!rd_l(f, &in->e_len)) return AF_ERR_IO;
  if (in->e_len < 1) { ufem_set_error("no element"); return AF_ERR_SML; }
  in->e_id = lalloc(in->e_len); in->e_type = lalloc(in->e_len); in->e_rs = lalloc(in->e_len);
  in->e_mp = lalloc(in->e_len); in->e_from = lalloc(in->e_len);
  if (!in->e_id || !in->e_type || !in->e_rs || !in->e_mp || !in->e_from) return AF_ERR_MEM;
  for (i = 0, sum = 0; i < in->e_len; i++) {
    if (!rd_l(f, &in->e_id[i]) || !rd_l(f, &in->e_type[i]) || !rd_l(f, &in->e_rs[i]) ||
        !rd_l(f, &in->e_mp[i]) || !rd_l(f, &c))
      return AF_ERR_IO;
    in->e_from[i] = sum;
    sum += c;
  }
  if (!rd_l(f, &in->e_len_l)) return AF_ERR_IO;
  if (in->e_len_l != sum) {
    ufem_set_error("element nodes: %ld given, %ld counted", in->e_len_l, sum);
    return AF_ERR_SIZ;
  }
  if (!(in->e_nodes = lalloc(in->e_len_l))) return AF_ERR_MEM;
  for (i = 0; i < in->e_len_l; i++)
    if (!rd_l(f, &in->e_nodes[i])) return AF_ERR_IO;
  /* gravity, nodal loads, element loads (fem/fem_io.c:602-606, :86-175) */
  if (!rd_l(f, &in->gr_dir) || !rd_d(f, &in->gr_val)) return AF_ERR_IO;
  if (!rd_l(f, &in->nl_len)) return AF_ERR_IO;
  if (in->nl_len < 0) in->nl_len = 0;
  in->nl_node = lalloc(in->nl_len); in->nl_type = lalloc(in->nl_len);
  in->nl_dir = lalloc(in->nl_len); in->nl_val = dalloc(in->nl_len);
  if (!in->nl_node || !in->nl_type || !in->nl_dir || !in->nl_val) return AF_ERR_MEM;
  for (i = 0; i < in->nl_len; i++)
    if (!rd_l(f, &in->nl_node[i]) || !rd_l(f, &in->nl_type[i]) || !rd_l(f, &in->nl_dir[i]) ||
        !rd_d(f, &in->nl_val[i]))
      return AF_ERR_IO;
  if (!rd_l(f, &in->el_len)) return AF_ERR_IO;
  if (in->el_len > 0) { ufem_set_error("element loads are not supported"); return AF_ERR_TYP; }
  /* array data and dynamics: a missing triple means none (fem/fem_io.c:185-193) */
  if (fscanf(f, "%li %li %lf", &trans_num, &dyn_num, &dyn_stp) == 3 && dyn_num > 0) {
    ufem_set_error("time-dependent data are not supported");
    return AF_ERR_TYP;
  }
  dyn0 = 0;
  if (fscanf(f, "%li", &dyn0) == 1 && dyn0 > 0) {
    ufem_set_error("dynamic loads are not supported");
    return AF_ERR_TYP;
  }
  /* Monte section: absent or ilen <= 0 means none (fem/fem_mont.c:872-874) */
  in->mc_fail_id = -1;
  {
    long il, ol;
    if (fscanf(f, "%li %li", &il, &ol) == 2 && il > 0) {
      long n = il + (ol > 0 ? ol : 0);
      in->mc_ilen = il;
      in->mc_olen = ol > 0 ? ol : 0;
      in->mc_pos = lalloc(n); in->mc_type = lalloc(n); in->mc_item = lalloc(n);
      in->mc_sub = lalloc(n); in->mc_iter = lalloc(n);
      if (!in->mc_pos || !in->mc_type || !in->mc_item || !in->mc_sub || !in->mc_iter) return AF_ERR_MEM;
      for (i = 0; i < n; i++) {
        if (!rd_l(f, &in->mc_pos[i]) || !rd_l(f, &in->mc_type[i]) || !rd_l(f, &in->mc_item[i]) ||
            !rd_l(f, &in->mc_sub[i]) || !rd_l(f, &in->mc_iter[i])) {
          ufem_set_error("incomplete random data");
          return AF_ERR_IO;
        }
        if (in->mc_type[i] == 19) in->mc_fail_id = i - il;   /* MONTE_VTYPE_RES_FAIL */
      }
    }
  }
  return AF_OK;
}

/* ---- IDs -> positions (the input form without -p) ------------------------
 * femRenumIndex (fem/fem_io.c:50-80) replaces every ID of a list by the
 * position of the first item carrying it, by a linear search of the item
 * IDs per entry (O(n m)).  Here the item IDs go into an open-addressing hash
 * table once (the first position of an ID kept), and every entry is one
 * lookup.  An ID no item carries is AF_ERR_VAL, as there. */
typedef struct {
  long *key, *val;
  unsigned long mask;
} idmap;

static unsigned long id_hash(long id)
{
  unsigned long z = (unsigned long)id + 0x9E3779B97F4A7C15UL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9UL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBUL;
  return z ^ (z >> 31);
}

static int idmap_build(idmap *m, const long *ids, long n)
{
  unsigned long cap = 16, h;
  long i;
  while (cap < 2UL * (unsigned long)n) cap <<= 1;
  m->key = (long *)malloc(sizeof(long) * cap);
  m->val = (long *)malloc(sizeof(long) * cap);
  m->mask = cap - 1;
  if (!m->key || !m->val) return AF_ERR_MEM;
  for (h = 0; h < cap; h++) m->val[h] = -1;
  for (i = 0; i < n; i++) {
    for (h = id_hash(ids[i]) & m->mask; m->val[h] >= 0 && m->key[h] != ids[i]; h = (h + 1) & m->mask) {}
    if (m->val[h] < 0) {      /* a repeated ID keeps its first position */
      m->key[h] = ids[i];
      m->val[h] = i;
    }
  }
  return AF_OK;
}

static void idmap_free(idmap *m)
{
  free(m->key);
  free(m->val);
}

static int renumber(long *ind, long n, const long *ids, long nids, const char *what)
{
  idmap m;
  long i;
  int rv = idmap_build(&m, ids, nids);
  if (rv != AF_OK) { idmap_free(&m); return rv; }
  for (i = 0; i < n; i++) {
    unsigned long h;
    for (h = id_hash(ind[i]) & m.mask; m.val[h] >= 0 && m.key[h] != ind[i]; h = (h + 1) & m.mask) {}
    if (m.val[h] < 0) {
      ufem_set_error("%s %ld: ID %ld names no item", what, i, ind[i]);
      idmap_free(&m);
      return AF_ERR_VAL;
    }
    ind[i] = m.val[h];
  }
  idmap_free(&m);
  return AF_OK;
}

/* the reference's order (fem/fem_io.c:568-605, :113): element material sets,
 * real sets (when there are any), element nodes, load nodes */
static int renumber_ids(ufem_input *in)
{
  int rv;
  if ((rv = renumber(in->e_mp, in->e_len, in->mp_id, in->mp_len, "element material set")) != AF_OK) return rv;
  if (in->rs_len > 0 && (rv = renumber(in->e_rs, in->e_len, in->rs_id, in->rs_len, "element real set")) != AF_OK)
    return rv;
  if ((rv = renumber(in->e_nodes, in->e_len_l, in->n_id, in->n_len, "element node")) != AF_OK) return rv;
  return renumber(in->nl_node, in->nl_len, in->n_id, in->n_len, "load node");
}

/* what the linear-static brick path supports: element type 9 with 8 nodes,
 * material type 1, no gravity, nodal loads of type 1 (displacement) and 3
 * (force) in directions 1-3, positions in range */
static int check_scope(ufem_input *in)
{
  long i, k;
  for (i = 0; i < in->mp_len; i++)
    if (in->mp_type[i] != 1) { ufem_set_error("material %ld: type %ld not supported", i, in->mp_type[i]); return AF_ERR_TYP; }
  for (i = 0; i < in->e_len; i++) {
    long cnt = (i + 1 < in->e_len ? in->e_from[i + 1] : in->e_len_l) - in->e_from[i];
    if (in->e_type[i] != 9 || cnt != 8) { ufem_set_error("element %ld: only 8-node bricks (type 9)", i); return AF_ERR_TYP; }
    if (in->e_mp[i] < 0 || in->e_mp[i] >= in->mp_len) { ufem_set_error("element %ld: material out of range", i); return AF_ERR_VAL; }
    for (k = 0; k < 8; k++) {
      long p = in->e_nodes[in->e_from[i] + k];
      if (p < 0 || p >= in->n_len) { ufem_set_error("element %ld: node out of range", i); return AF_ERR_VAL; }
    }
  }
  if (in->gr_val != 0.0) { ufem_set_error("gravity loads are not supported"); return AF_ERR_TYP; }
  for (i = 0; i < in->nl_len; i++) {
    if (in->nl_type[i] != 1 && in->nl_type[i] != 3) { ufem_set_error("load %ld: type %ld not supported", i, in->nl_type[i]); return AF_ERR_TYP; }
    if (in->nl_node[i] < 0 || in->nl_node[i] >= in->n_len || in->nl_dir[i] < 1 || in->nl_dir[i] > 3) {
      ufem_set_error("load %ld out of range", i);
      return AF_ERR_VAL;
    }
  }
  return AF_OK;
}

int ufem_input_read_form(const char *fname, int preparsed, ufem_input *in)
{
  FILE *f;
  int rv;
  memset(in, 0, sizeof(*in));
  in->mc_fail_id = -1;
  if (!fname || !(f = fopen(fname, "r"))) { ufem_set_error("cannot open %s", fname ? fname : "(null)"); return AF_ERR_IO; }
  rv = read_body(f, in);
  fclose(f);
  if (rv == AF_ERR_IO) ufem_set_error("%s: short or unreadable input", fname);
  if (rv == AF_OK && !preparsed) rv = renumber_ids(in);
  if (rv == AF_OK) rv = check_scope(in);
  if (rv != AF_OK) ufem_input_free(in);
  return rv;
}

int ufem_input_read(const char *fname, ufem_input *in)
{
  return ufem_input_read_form(fname, 1, in);
}

/* ---- the public model handle -------------------------------------------- */
struct ufem_model {
  ufem_input in;
  double *E, *nu;
};

ufem_model *ufem_model_read(const char *fname, int preparsed, int *rv)
{
  ufem_model *m = (ufem_model *)calloc(1, sizeof(ufem_model));
  int r;
  if (!m) { if (rv) *rv = AF_ERR_MEM; return NULL; }
  if ((r = ufem_input_read_form(fname, preparsed, &m->in)) != AF_OK) {
    free(m);
    if (rv) *rv = r;
    return NULL;
  }
  m->E = dalloc(m->in.e_len);
  m->nu = dalloc(m->in.e_len);
  if (!m->E || !m->nu) {
    ufem_model_free(m);
    if (rv) *rv = AF_ERR_MEM;
    return NULL;
  }
  ufem_input_elastic(&m->in, m->E, m->nu);
  if (rv) *rv = AF_OK;
  return m;
}

void ufem_model_free(ufem_model *m)
{
  if (!m) return;
  ufem_input_free(&m->in);
  free(m->E);
  free(m->nu);
  free(m);
}

int ufem_model_mesh(const ufem_model *m, ufem_mesh *mesh)
{
  if (!m || !mesh) return AF_ERR_VAL;
  memset(mesh, 0, sizeof(*mesh));
  mesh->n_nodes = m->in.n_len;
  mesh->x = m->in.n_x;
  mesh->y = m->in.n_y;
  mesh->z = m->in.n_z;
  mesh->n_elems = m->in.e_len;
  mesh->enodes = m->in.e_nodes;      /* 8 per element, in order (checked on reading) */
  mesh->E = m->E;
  mesh->nu = m->nu;
  mesh->n_nl = m->in.nl_len;
  mesh->nl_node = m->in.nl_node;
  mesh->nl_type = m->in.nl_type;
  mesh->nl_dir = m->in.nl_dir;
  mesh->nl_val = m->in.nl_val;
  return AF_OK;
}

const long *ufem_model_node_ids(const ufem_model *m) { return m ? m->in.n_id : NULL; }
const long *ufem_model_elem_ids(const ufem_model *m) { return m ? m->in.e_id : NULL; }
const long *ufem_model_elem_mat(const ufem_model *m) { return m ? m->in.e_mp : NULL; }
const long *ufem_model_elem_rset(const ufem_model *m) { return m ? m->in.e_rs : NULL; }
long ufem_model_sol_id(const ufem_model *m) { return m ? m->in.sol_id : 0; }
long ufem_model_sol_dep(const ufem_model *m) { return m ? m->in.sol_dep_on : 0; }

/* femGetRepValIndex (fem/fem_accs.c:297-381) over material type 1: a
 * non-repeating value missing from the list or past the set gives -1; a
 * repeating one past the set gives -1 and one missing from the list gives 0,
 * both as the reference returns them */
long ufem_input_mat_index(const ufem_input *in, long mp, long type, long rep)
{
  long i, i_pos = -1, len;
  if (mp < 0 || mp >= in->mp_len) return -1;
  len = (mp + 1 >= in->mp_len ? in->mp_len_l : in->mp_from[mp + 1]) - in->mp_from[mp];
  if (rep <= 0) {
    for (i = 0; i < 7; i++)
      if (mat1_val[i] == type) { i_pos = i; break; }
    if (i_pos < 0) return -1;
  } else {
    for (i = 0; i < 6; i++)
      if (mat1_val_rp[i] == type) { i_pos = i + 6 * (rep - 1) + 7; break; }
    if (i_pos < 0) return 0;
  }
  if (i_pos >= len) return -1;
  return in->mp_from[mp] + i_pos;
}

/* femGetMPValPos(e, MAT_EX / MAT_NU, 0) (fem/fem_dama.c:410-436): 0 when the
 * set lacks the value */
void ufem_input_elastic(const ufem_input *in, double *E, double *nu)
{
  long e, k;
  for (e = 0; e < in->e_len; e++) {
    k = ufem_input_mat_index(in, in->e_mp[e], MAT_EX, 0);
    E[e] = k >= 0 ? in->mp_val[k] : 0.0;
    k = ufem_input_mat_index(in, in->e_mp[e], MAT_NU, 0);
    nu[e] = k >= 0 ? in->mp_val[k] : 0.0;
  }
}
