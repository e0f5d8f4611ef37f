/*
 * ugk.h -- thin C-ABI between the host C code of libufem_gpu.so and its
 * hand-written HIP kernels for gfx950 (ugk_*.hip).  Plain pointers and sizes
 * only; every pointer named d_* is device memory.
 *
 * Matrix storage on the device is "SELL-64 of b x b blocks": block rows are
 * grouped into slices of 64 (one wavefront, lane = block row); slice s owns
 * slen[s] slots; the column of slot k of lane l is cols[(soff[s] + k) * 64 + l]
 * and its b*b values sit at vals[((soff[s] + k) * b*b + q) * 64 + l] for
 * q = r*b + c, so every load of a wavefront is 64 consecutive words.  Slots
 * past a row's length are padding: column = the row itself, value 0.
 * b = 3 for meshes (node blocks), b = 1 for matrices handed over in the
 * reference's scalar row-slot form.
 */
#ifndef UGK_H
#define UGK_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int bs;             /* block size b: 1 or 3                           */
  long n_brows;       /* block rows                                     */
  long n_slices;      /* ceil(n_brows / 64)                             */
  long total_slots;   /* sum of slen                                    */
  long nnzb;          /* real (unpadded) blocks                         */
  int *d_slen;        /* [n_slices]                                     */
  long *d_soff;       /* [n_slices] slot offset of each slice           */
  int *d_rowlen;      /* [n_brows] real blocks per row                  */
  int *d_cols;        /* [total_slots * 64]                             */
  double *d_vals;     /* [total_slots * 64 * b*b]                       */
  float *d_vals32;    /* FP32 copy for mixed precision (may be NULL); b = 3:
                         each lane's first eight values of a slot in two
                         runs of four (256 k + 4 l), the ninth at 512 + l */
  unsigned short *d_rem16;  /* bf16 of vals - vals32 at vals32's positions:
                         with it the copy holds 48-bit values (products
                         with FP64 vectors only; may be NULL)           */
  float *d_corr32;    /* b = 3: [9][n_brows] symmetric part of the block row
                         sums of (vals - (vals32 + rem16)), added to the
                         FP32 copy's diagonal blocks (may be NULL)      */
  unsigned long gen;  /* bumped whenever the pattern is (re)uploaded     */
} ugk_mat;

/* position of value q (= 3 r + c, row-major in the block) of lane l in slot S
 * of a 3x3-block value array: a slot holds 576 values, each lane's first eight
 * in adjacent pairs (pair k of lane l at 128 k + 2 l: one 16-byte load in
 * FP64), the ninth in a 64-wide tail at 512 + l.  1x1 blocks: 64 S + l. */
static inline long ugk_vpos3(long S, int q, int l)
{
  return S * 576 + (q < 8 ? (q >> 1) * 128 + 2 * l + (q & 1) : 512 + l);
}

/* largest grid of a reducing kernel, and the size (in doubles) of the
 * partial-sum buffer every reducing call takes (8 slots per block + counter) */
#define UGK_RED_MAX_BLOCKS 65536
#define UGK_PART_DOUBLES (8L * UGK_RED_MAX_BLOCKS + 16)

/* solver scalars kept on the device between kernels */
typedef struct {
  double rz[2];       /* (r, z) of the current / next iteration         */
  double alpha, beta, omega, rho;
  double rr, bb, xx;  /* last ||r||^2, ||b||^2, ||x||^2                 */
  double normA;       /* max row 2-norm (reference stopping rule)       */
  double tol;
  double aux[8];
  long iter;          /* completed iterations                           */
  double rzB;         /* (r, z) of the running iteration, for its update */
  long pidx;          /* which ping-pong buffer holds the current p      */
  double maxrr[2];    /* mixed: largest sloppy ||r|| since the last true residual */
  double rr_true;     /* mixed: ||b - K x||^2 of the last FP64 residual   */
  double delta;       /* mixed: refresh when the sloppy ||r|| drops by this */
  long ru_req, ru_at; /* mixed: iteration of the last refresh request / done */
  long n_ru;          /* mixed: FP64 residual refreshes                   */
  long cur;           /* single-reduction PCG: iteration latched by its SpMV */
  double gam[2], alf[2];  /* single-reduction PCG: gamma, alpha by parity  */
  double best_true;   /* mixed: smallest true ||r||^2 seen                */
  long n_stall;       /* mixed: refreshes in a row without progress       */
  long max_gap;       /* mixed: refresh at least every max_gap iterations */
  long maxit;
  int rule;           /* 0 reference rule, 1 relative residual          */
  int done;           /* 0 running, 1 converged, 2 breakdown, 3 maxit   */
  int err;            /* kernel-detected error (zero diagonal, detJ<=0)  */
  int restart;        /* mixed: restart the inner CG at every FP64 refresh */
  double net_s;       /* flexible CG: rms of sqrt(dinv) r (network input scale) */
  double pqs[8];      /* flexible CG: (p, A p) of the kept directions     */
  long fb_iter;       /* flexible CG: iteration of the Jacobi fallback    */
  int fb;             /* flexible CG: 1 once the network is switched off  */
  int halt;           /* flexible CG: done or fallback (network skipped)  */
  /* flexible CG, one kept direction (ugk_f1_*) */
  long hgrid;         /* grid of the kernel that wrote the head partials   */
  long restart_at;    /* iteration whose direction starts afresh (beta 0)  */
  double rdr;         /* (r, dinv r) of the running iteration              */
  double pr_old;      /* (p_{k-1}, r_k)                                    */
  double pq_old;      /* (p_{k-1}, q_{k-1}), q the stored product           */
  long zgrid;         /* grid of the combine that wrote the z partials     */
  unsigned int counter[10]; /* last-block tickets, reset by the last block;
                               6..8: the persistent kernel's grid barrier */
} ugk_state;

/* ---------------- fixed workspace for callers that cannot allocate -------- */
/* the device address (for the current device) and size of a file-scope
 * workspace of 6 vectors of UGK_NOALLOC_MAX_N doubles, UGK_PART_DOUBLES
 * reduction partials and UGK_NOALLOC_STATE_DOUBLES for a ugk_state, in that
 * order; taken with hipGetSymbolAddress once per device */
#define UGK_NOALLOC_MAX_N (1L << 20)
#define UGK_NOALLOC_STATE_DOUBLES 128
int ugk_noalloc_workspace(void **ptr, size_t *bytes);

/* ---------------- memory / stream / timing ------------------------------ */
int   ugk_init(void);
const char *ugk_error(void);
void *ugk_malloc(size_t bytes);                  /* zero-filled; NULL on failure */
void  ugk_free(void *p);
int   ugk_h2d(void *d, const void *h, size_t bytes, void *stream);
int   ugk_d2h(void *h, const void *d, size_t bytes, void *stream);
int   ugk_d2d(void *d, const void *s, size_t bytes, void *stream);
int   ugk_zero(void *d, size_t bytes, void *stream);  /* fill kernel, graph-safe */
void *ugk_stream_create(void);
void  ugk_stream_destroy(void *stream);
int   ugk_sync(void *stream);
void *ugk_event_create(void);
void  ugk_event_destroy(void *ev);
int   ugk_event_record(void *ev, void *stream);
float ugk_event_ms(void *ev0, void *ev1);        /* after both completed */
int   ugk_device_info(int *n_cu, long *hbm_bytes, char *name, int name_len);
void *ugk_host_alloc(size_t bytes);              /* pinned host memory */
/* pinned, coherent host memory the device reads while a kernel runs; *dptr
 * gets its device address (free with ugk_host_free) */
void *ugk_host_alloc_mapped(size_t bytes, void **dptr);
/* 1 once the event's work is done, 0 while pending, -1 on error */
int   ugk_event_done(void *ev);
void  ugk_host_free(void *p);
int   ugk_d2h_async(void *h, const void *d, size_t bytes, void *stream);
int   ugk_event_sync(void *ev);

/* graphs: capture what is enqueued on `stream` between begin and end */
int   ugk_graph_begin(void *stream);
void *ugk_graph_end(void *stream);               /* returns executable graph */
int   ugk_graph_launch(void *graph, void *stream);
void  ugk_graph_destroy(void *graph);

/* ---------------- assembly ---------------------------------------------- */
/* element stiffness, one wavefront per element: dense 24x24 per element into
 * d_Ke (row-major), or added into the SELL-3 matrix for the elements listed in
 * d_elist (one colour: no two share a node).  npts = 2 or 3 Gauss points per
 * axis.  Sets st->err = 1 on a non-positive Jacobian. */
/* strains and stresses at the npts^3 Gauss points of every element of a
 * brick mesh for the displacement vector d_u (3 per node): d_out
 * [ne][npts^3][12], sigma x, y, z, yz, zx, xy then epsilon (engineering
 * shears), point (i npts + j) npts + k along xi, eta, zeta */
int ugk_elem_results(long ne, int npts, const double *d_x, const double *d_y, const double *d_z,
                     const int *d_enodes, const double *d_E, const double *d_nu, const double *d_u, double *d_out,
                     ugk_state *d_st, void *stream);
int ugk_elem_stiff_dense(long n_elems, const double *d_xyz, const int *d_enodes,
                         const double *d_E, const double *d_nu, int npts,
                         double *d_Ke, ugk_state *d_st, void *stream);
int ugk_elem_stiff_scatter(const ugk_mat *K, long n_list, const int *d_elist,
                           const double *d_x, const double *d_y, const double *d_z,
                           const int *d_enodes, const double *d_E, const double *d_nu,
                           int npts, ugk_state *d_st, void *stream);
/* supports: copy the rows of the fixed DOFs (for reactions), then zero those
 * rows and columns and put 1 on their diagonal */
int ugk_save_rows(const ugk_mat *K, long n_fixed, const int *d_fixed, int maxb,
                  double *d_rows, int *d_rcols, void *stream);
int ugk_apply_zero_bc(const ugk_mat *K, long n_fixed, const int *d_fixed, void *stream);
/* R[k] = (K_orig u)[fixed[k]] - F_orig[fixed[k]] from the saved rows */
int ugk_reactions(const ugk_mat *K, long n_fixed, const int *d_fixed, int maxb,
                  const double *d_rows, const int *d_rcols, const double *d_u,
                  const double *d_Forig, double *d_R, void *stream);
/* FP32 copy of the values */
int ugk_vals_to_f32(const ugk_mat *K, void *stream);
/* d_dst[d_map[i]] = d_src[i] for i < n (a matrix's values from host order) */
int ugk_scatter_vals(long n, const int *d_map, const double *d_src, double *d_dst, void *stream);

/* ---------------- sparse kernels ---------------------------------------- */
int ugk_spmv(const ugk_mat *K, const double *d_x, double *d_y, void *stream);
/* streaming read of K's values and columns (plain or non-temporal loads);
 * one partial per wave into d_out */
int ugk_stream_read(const ugk_mat *K, int nontemporal, double *d_out, void *stream);
int ugk_spmv_f32(const ugk_mat *K, const float *d_x, float *d_y, void *stream);
/* y = K x with the FP32 copy of K (and its diagonal correction), FP64 vectors */
int ugk_spmv_f32x64(const ugk_mat *K, const double *d_x, double *d_y, void *stream);
/* dinv[i] = 1 / K_ii; st->err = 2 if |K_ii| < 1e-17 */
int ugk_diag_inv(const ugk_mat *K, double *d_dinv, ugk_state *d_st, void *stream);
/* max row 2-norm into st->normA */
int ugk_mat_norm(const ugk_mat *K, ugk_state *d_st, double *d_part, void *stream);

/* ---------------- vector kernels (FP64) --------------------------------- */
int ugk_dot(long n, const double *d_a, const double *d_b, double *d_out, double *d_part, void *stream);
int ugk_axpby(long n, double a, const double *d_x, double b, const double *d_y, double *d_z, void *stream);

/* ---------------- PCG (Jacobi), streamed form ---------------------------- */
/* start: r = b - A x, z = dinv r, p0 = p1 = 0; block partials of (r,z),
 * (r,r), (x,x), (b,b) for the first iteration kernel; st->iter must be 0 */
int ugk_pcg_start(const ugk_mat *K, long n, const double *d_b, const double *d_x,
                  double *d_r, double *d_z, double *d_p0, double *d_p1, const double *d_dinv,
                  ugk_state *d_st, double *d_part, void *stream);
/* one iteration; no-ops once st->done != 0.  variant 0: two kernels (the
 * p update folded into the SpMV gather); variant 1: three kernels (p update,
 * SpMV, vector update).  Iteration state lives on the device only. */
int ugk_pcg_iter(const ugk_mat *K, long n, double *d_x, double *d_r, double *d_z,
                 double **d_p /* [2] ping-pong */, double *d_q, const double *d_dinv,
                 ugk_state *d_st, double *d_part, int variant, void *stream);
/* one kernel of the iteration alone, for timing: the fused p-update + SpMV
 * (ugk_pcg_iter_a) or the streamed form's q = A p + (p, q) (ugk_pcg_iter_q) */
int ugk_pcg_iter_a(const ugk_mat *K, const double *d_z, const double *d_pold, double *d_pnew,
                   double *d_q, ugk_state *d_st, double *d_part, void *stream);
int ugk_pcg_iter_q(const ugk_mat *K, const double *d_p, double *d_q, ugk_state *d_st, double *d_part,
                   void *stream);
int ugk_cg1_iter_w(const ugk_mat *K, const double *d_u, double *d_w, ugk_state *d_st, double *d_part,
                   void *stream);
/* single-reduction (Chronopoulos-Gear) Jacobi PCG: start, then per iteration
 * w = A u with (w,u) partials, and one vector kernel doing the reduction,
 * the convergence test and p, s, x, r, u updates.  d_s, d_w: two more vectors */
int ugk_cg1_start(const ugk_mat *K, long n, const double *d_b, const double *d_x, double *d_r,
                  double *d_u, double *d_p, double *d_s, const double *d_dinv, double *d_part,
                  void *stream);
int ugk_cg1_iter(const ugk_mat *K, long n, double *d_x, double *d_r, double *d_u, double *d_p,
                 double *d_s, double *d_w, const double *d_dinv, ugk_state *d_st, double *d_part,
                 void *stream);
/* whole solve in one cooperative launch (grid barriers between phases) */
int ugk_pcg_persistent(const ugk_mat *K, long n, double *d_x, double *d_r, double *d_z,
                       double *d_p0, double *d_p1, double *d_q, const double *d_dinv,
                       ugk_state *d_st, double *d_part, void *stream);

/* ---------------- mixed precision: FP32 CG with FP64 refreshes ----------- */
/* start: r64 = b - K x64 (FP64), r32 = r64, z32 = dinv32 r32, p32 = x32 = 0 */
int ugk_mx_start(const ugk_mat *K, long n, const double *d_b, const double *d_x64, double *d_r64,
                 float *d_r32, float *d_z32, float *d_p32, float *d_x32, const float *d_dinv32,
                 double *d_part, void *stream);
/* one iteration: p update, q = K32 p (FP32), x32/r32/z32 update, refresh test,
 * and -- when the FP32 residual has dropped by delta -- x64 += x32 and
 * r64 = b - K x64 in FP64 */
int ugk_mx_iter(const ugk_mat *K, long n, const double *d_b, double *d_x64, double *d_r64,
                float *d_r32, float *d_z32, float *d_p32, float *d_q32, float *d_x32,
                const float *d_dinv32, ugk_state *d_st, double *d_part, void *stream);
int ugk_to_f32(long n, const double *d_a, float *d_b, void *stream);
/* FP64 residual r = b - K x with ||r||^2 into st->rr (one-off, not in loops) */
int ugk_residual(const ugk_mat *K, long n, const double *d_b, const double *d_x, double *d_r,
                 ugk_state *d_st, double *d_part, void *stream);

/* ---------------- BiCGStab (Jacobi, right preconditioned) ---------------- */
int ugk_bicgs_start(const ugk_mat *K, long n, const double *d_b, const double *d_x,
                    double *d_r, double *d_rhat, double *d_p, double *d_v,
                    ugk_state *d_st, double *d_part, void *stream);
int ugk_bicgs_iter(const ugk_mat *K, long n, double *d_x, double *d_r, const double *d_rhat,
                   double *d_p, double *d_phat, double *d_v, double *d_s, double *d_shat,
                   double *d_t, const double *d_dinv, ugk_state *d_st, double *d_part,
                   void *stream);

/* ---------------- batched solve: one persistent launch ------------------- */
/* one sample: its K (pattern shared within a mesh class, values own), Jacobi
 * diagonal inverse, right-hand side and Krylov vectors (device pointers) */
typedef struct {
  const int *slen;
  const long *soff;
  const int *cols;
  long n_brows, n_slices;
  const double *vals;
  const double *dinv;
  const double *f;
  double *x, *r, *u, *p, *s, *w;
  long maxit;         /* per-sample iteration cap (0: the launch's) */
  const float *vals32;  /* mixed: the FP32 copy of vals (NULL: FP64 products) */
  const float *corr32;  /* mixed: its row-sum correction [9][n_brows] */
  const unsigned short *rem16;  /* mixed, may be NULL: bf16 of vals - vals32 */
} ugk_bsample;

typedef struct {
  long iters;
  double relres;      /* recursive ||r|| / ||f|| at exit */
  int status;         /* 1 converged, 2 breakdown, 3 maxit */
  int team;           /* workgroups that solved it */
  long t_start, t_end;   /* device wall clock ticks */
} ugk_bresult;

/* Every workgroup repeatedly takes the next slot (a ticket counter over
 * d_slots) and works on that slot's sample as one rank of a team: slot =
 * {sample, (team size << 8) | rank}, a team's slots adjacent.  Jacobi PCG
 * (single reduction) with workgroup barriers, and a team barrier on
 * d_tctr[sample] when the team has more than one workgroup (d_tpart:
 * 64 doubles per sample).  d_tctr must be zero at launch.  The grid must not
 * exceed the residency (ugk_batch_occupancy), so that every team forms.
 * wg: 256, 512 or 1024 threads.  A team wait gives up after wait_ticks of
 * the device wall clock (<= 0: no practical limit) or when the int at
 * d_abort (mapped host memory, may be NULL) becomes nonzero; the sample then
 * reports status 4, and no workgroup takes a new sample after a stop.
 * d_ticket holds 2 zeroed words (ticket, device stop flag).  A sample with
 * vals32 takes w = K u from the FP32 copy (plus rem16 when set, and its
 * correction), every
 * vector in FP64, and replaces r by f - K x in FP64 (keeping the direction:
 * reliable update) each time ||r|| has fallen by 10 since the last such
 * refresh or after 500 iterations; it converges only on a refreshed
 * residual, and a breakdown of the recurrence (p, K p) <= 0 restarts it
 * along dinv r from a refreshed residual before it counts as status 2. */
int ugk_batch_pcg(const ugk_bsample *d_samples, const int *d_slots, int n_slots, unsigned *d_ticket,
                  unsigned *d_tctr, double *d_tpart, ugk_bresult *d_res, double tol, long maxit,
                  int wg, int grid, const int *d_abort, long wait_ticks, void *stream);
/* blocks of `wg` threads the batch kernel can keep resident per CU */
int ugk_batch_occupancy(int wg);
/* dinv = 1 / diag(K) for one block-3 matrix (no state) */
int ugk_diag_inv3(const ugk_mat *K, double *d_dinv, ugk_state *d_st, void *stream);
long ugk_wall_clock_khz(void);

/* ---------------- flexible CG with a nonlinear preconditioner ------------- */
/* FCG(m): directions are A-orthogonalised against the last m (1..8), kept in
 * ring slots (iteration mod m) of d_P and d_Q (m vectors of n each).  Per
 * iteration: ugk_fcg_head (convergence test on the previous update's
 * partials; network input X0[v][c] = sqrt(dinv) r / s, s = rms(sqrt(dinv) r),
 * for c < 3 in the split format [V][3 parts][4], channel 3 left alone), the network (ugk_conv and friends,
 * skipped through st->halt), ugk_fcg_combine (z = dinv r + s mask sqrt(dinv)
 * out and the dots with the kept directions), ugk_fcg_p (coefficients, descent check and
 * Jacobi fallback, new direction), ugk_fcg_q (q = A p), ugk_fcg_update.
 * out: f32 [n/3][4].
 * Mixed form (state max_gap > 0, delta, ru_req = -1, ru_at = 0): ugk_fcg_q
 * with f32 = 1 multiplies by the FP32 copy of A; ugk_fcg_refresh, launched
 * before each head, recomputes r = b - A x in FP64 for the iterations the
 * head asked for (the FP32-based residual has dropped by delta since the
 * last refresh, max_gap iterations have passed, or it passed the test: a
 * convergence is only accepted on a recomputed residual). */
int ugk_fcg_start(const ugk_mat *K, long n, const double *d_b, const double *d_x, double *d_r,
                  const double *d_dinv, double *d_part, void *stream);
int ugk_fcg_head(const ugk_mat *K, long n, const double *d_r, const double *d_dinv, unsigned short *d_x0,
                 ugk_state *d_st, double *d_part, void *stream);
int ugk_fcg_combine(long n, const double *d_r, const double *d_dinv, const float *d_mask,
                    const float *d_out, const double *d_P, const double *d_Q, int m, const double *d_zc,
                    double *d_z, ugk_state *d_st, double *d_part, void *stream);
int ugk_fcg_p(long n, const double *d_r, const double *d_dinv, const double *d_z, double *d_P, int m,
              ugk_state *d_st, double *d_part, void *stream);
int ugk_fcg_q(const ugk_mat *K, long n, const double *d_P, double *d_Q, int m, int f32, ugk_state *d_st,
              double *d_part, void *stream);
int ugk_fcg_refresh(const ugk_mat *K, const double *d_b, const double *d_x, double *d_r, const double *d_dinv,
                    ugk_state *d_st, double *d_part, void *stream);
int ugk_fcg_update(const ugk_mat *K, long n, double *d_x, double *d_r, const double *d_P, const double *d_Q,
                   int m, const double *d_dinv, ugk_state *d_st, double *d_part, void *stream);

struct ugk_conv_desc_s;
/* FCG(1) in three kernels per iteration (one kept direction; the network or
 * the coarse term, or both, in z = M r).  Per iteration k:
 *   [ugk_f1_refresh]  mixed form only: r = b - A x in FP64 when asked for;
 *   the network on the input the previous update (or refresh) wrote;
 *   [ugk_cc_apply]    the coarse term into d_zc;
 *   ugk_f1_combine    convergence test of r_k (and, mixed, the refresh
 *                     request), z = dinv r + s mask sqrt(dinv) out (+ zc),
 *                     partials (z, r), (z, q_{k-1});
 *   ugk_f1_w          w = A z (FP64, or the FP32 copy with f32 = 1), (z, w);
 *   ugk_f1_update     beta = -(z, q_{k-1}) / (p_{k-1}, q_{k-1}),
 *                     p = z + beta p, q = w + beta q (= A p), alpha =
 *                     (p, r) / (p, q) with (p, q) from the symmetric
 *                     identity, x and r update, the next network input
 *                     d_x0 (split format, scaled by the rms of the last r) and the
 *                     head partials of r_{k+1}.
 * A direction that does not descend (or is not finite) is not taken: that
 * iteration moves nothing and the next one starts afresh with z = dinv r,
 * the network skipped from there on (st->fb, st->halt).  ugk_f1_start forms r = b - A x and its
 * partials; ugk_f1_input0 writes the first network input.  The network input
 * (written by input0, refresh and update) is d_x0, split f32 of
 * sqrt(dinv) r / s, or, when d_v64 is given, FP64 mask sqrt(dinv) r
 * (unscaled: the FP64 inference path).  A refresh keeps
 * the direction (st->restart = 0) or starts it afresh (1). */
int ugk_f1_start(const ugk_mat *K, long n, const double *d_b, const double *d_x, double *d_r, const double *d_dinv,
                 ugk_state *d_st, double *d_part, void *stream);
int ugk_f1_input0(long n, const double *d_r, const double *d_dinv, unsigned short *d_x0, double *d_v64,
                  const float *d_mask, ugk_state *d_st, double *d_part, void *stream);
int ugk_f1_refresh(const ugk_mat *K, long n, const double *d_b, const double *d_x, double *d_r, const double *d_dinv,
                   const double *d_p, const double *d_q, unsigned short *d_x0, double *d_v64, const float *d_mask,
                   ugk_state *d_st, double *d_part, void *stream);
int ugk_f1_combine(long n, const double *d_r, const double *d_dinv, const float *d_mask, const float *d_out,
                   const double *d_zc, const double *d_q, double *d_z, ugk_state *d_st, double *d_part, void *stream);
int ugk_f1_w(const ugk_mat *K, const double *d_z, double *d_w, int f32, const ugk_state *d_st, double *d_part,
             void *stream);
/* ugk_f1_combine with the network's last layer folded in: y = C0^T t0, the
 * stride-2 transpose (UGK_CONV_T2) described by *cd (16 or 32 input
 * channels, 3 outputs, one sample) on the vector ALUs: f32 for an f32 input
 * (scaled by st->net_s), FP64 for UGK_ACT_F64 (unscaled); the network then
 * runs all stages but that one */
int ugk_f1_combine_c0t(long n, const double *d_r, const double *d_dinv, const float *d_mask,
                       const struct ugk_conv_desc_s *cd, const double *d_zc, const double *d_q, double *d_z,
                       ugk_state *d_st, double *d_part, void *stream);
int ugk_f1_update(const ugk_mat *K, long n, double *d_x, double *d_r, double *d_p, double *d_q, const double *d_z,
                  const double *d_w, const double *d_dinv, unsigned short *d_x0, double *d_v64, const float *d_mask,
                  ugk_state *d_st, double *d_part, void *stream);

/* ---------------- CG with the SSOR preconditioner ------------------------ */
/* Kc: A with block rows (and columns) in colour order, new row I' = old row
 * perm[I']; colour c owns rows [cb[c], cb[c+1]) and no two of its rows
 * couple.  ugk_ssor_permute_vals fills Kc's values from A (pattern already
 * on the device); ugk_ssor_permute_vec moves vectors into (dir 0) and out
 * of (dir 1) colour order.  One ugk_ssor_iter = convergence test, forward
 * and backward symmetric Gauss-Seidel sweeps (one launch per colour),
 * (r, z), p = z + beta p, q = Kc p, x and r update; all vectors in colour
 * order, r and the state's partials made by ugk_fcg_start on Kc. */
int ugk_ssor_permute_vals(const ugk_mat *K, const ugk_mat *Kc, const int *d_perm, void *stream);
int ugk_ssor_permute_vec(long n_brows, int bs, const int *d_perm, const double *d_src, double *d_dst, int dir,
                         void *stream);
int ugk_ssor_iter(const ugk_mat *Kc, int ncol, const long *cb, long n, double *d_x, double *d_r, double *d_zz,
                  double *d_z, double *d_p, double *d_q, const double *d_dinv, ugk_state *d_st, double *d_part,
                  void *stream);

/* The same solve in Eisenstat's form (see ugk_kernels.hip): CG on
 * D^1/2 W^-1 Kc W^-T D^1/2 (W = D + L in colour order) for the correction
 * x^ of x = x0 + W^-T D^1/2 x^, one backward and one forward sweep per
 * iteration over the slots above / below each row's diagonal (dpos).
 * start: sqrt of the diagonal into dsq, r^ = p^ = D^1/2 W^-1 r0 from
 * r0 = b - Kc x0 (r0 may be r^ itself), x^ = 0.  iter: one iteration; with
 * check, the original residual r = W D^-1/2 r^ (and for rule 0 the current
 * x into yx's sweep) is tested.  finish: x = x0 + W^-T D^1/2 x^ in place of
 * x0.  Vectors in colour order; sv, t, qh and yx are work vectors. */
int ugk_eis_check_size(const ugk_mat *Kc, int ncol);
int ugk_eis_start(const ugk_mat *Kc, const int *d_dpos, int ncol, const long *cb, long n, const double *d_r0,
                  double *d_rh, double *d_ph, double *d_xh, double *d_sv, double *d_dsq, ugk_state *d_st,
                  double *d_part, void *stream);
int ugk_eis_iter(const ugk_mat *Kc, const int *d_dpos, int ncol, const long *cb, long n, double *d_xh,
                 double *d_rh, double *d_ph, double *d_qh, double *d_t, double *d_sv, const double *d_dsq,
                 const double *d_x0, double *d_yx, int check, int rule, ugk_state *d_st, double *d_part,
                 void *stream);
int ugk_eis_finish(const ugk_mat *Kc, const int *d_dpos, int ncol, const long *cb, const double *d_xh,
                   double *d_x, double *d_yx, ugk_state *d_st, double *d_part, void *stream);

/* ---------------- two-level (coarse-grid) correction ---------------------- */
/* On a box node grid W x H x D (node x + W (y + H z), element
 * ex + (W-1) (ey + (H-1) ez)), coarsened by cx, cy, cz into a grid of
 * Wc x Hc x Dc nodes (W - 1 = cx (Wc - 1), ...).  P interpolates trilinearly
 * from the coarse nodes (cells of cx x cy x cz elements) to the fine nodes;
 * fm masks the fixed fine DOFs, cm the coarse DOFs whose coincident fine DOF
 * is fixed (their rows and columns of A are the identity).  The coarse
 * matrix is the Galerkin product A = P^T K P, assembled per cell from the
 * elements' Lame constants: K_cell = sum_p lambda_p G_p^l + mu_p G_p^m with
 * G_p = T_p^T K^(l|m) T_p of the cell's element p (identical box elements),
 * then inverted in place (dense, block Gauss-Jordan).  The correction is
 * z_c = fm P A^-1 P^T fm r. */
#define UGK_CC_MAX_CELL 4096   /* bricks per coarse cell */
typedef struct {
  int W, H, D, cx, cy, cz, Wc, Hc, Dc;
  long ncell;               /* (Wc - 1) (Hc - 1) (Dc - 1)                    */
  long nc;                  /* coarse DOFs, 3 Wc Hc Dc                       */
  long ncp;                 /* nc padded to a multiple of 64                 */
  double *d_G;              /* [cx cy cz][2][24 x 24]                        */
  double *d_KC;             /* [ncell][24 x 24]                              */
  double *d_A;              /* [ncp][ncp]: A, then A^-1                      */
  double *d_piv, *d_R, *d_C;/* block Gauss-Jordan: [B][B], [B][ncp], [ncp][B], B = UGK_GJ_BLOCK */
  double *d_cpart;          /* [24][ncell] restriction partials              */
  double *d_g, *d_y;        /* [ncp] coarse vectors                          */
  double *d_fm;             /* [3 W H D] fine mask                           */
  double *d_zc;             /* [3 W H D] the correction, for flexible CG     */
  double *d_cm;             /* [nc] coarse mask                              */
  const double *d_E, *d_nu; /* the system's element moduli                   */
} ugk_coarse;
/* K_cell from E, nu; dense A (identity at masked DOFs and in the padding) */
int ugk_cc_galerkin(ugk_coarse *C, void *stream);
int ugk_cc_assemble(ugk_coarse *C, void *stream);
/* in-place inverse of an SPD n x n matrix (n a multiple of 64, leading
 * dimension n) by block Gauss-Jordan without pivoting, blocks of
 * UGK_GJ_BLOCK; work buffers piv [B x B], R [B x n], Cb [n x B] */
#define UGK_GJ_BLOCK 64
int ugk_dense_spd_inverse(double *d_A, long n, double *d_piv, double *d_R, double *d_Cb, void *stream);
/* PCG with z = dinv r + z_c: convergence test, restriction, coarse solve,
 * z and (r, z), p = z + beta p, q = K p, x and r update (r and the
 * partials from ugk_fcg_start) */
int ugk_cc_iter(const ugk_mat *K, const ugk_coarse *C, long n, double *d_x, double *d_r, double *d_z,
                double *d_p, double *d_q, const double *d_dinv, ugk_state *d_st, double *d_part, void *stream);
/* the same with q = K p from the FP32 copy of K and the FP64 residual
 * b - K x recomputed when the flexible-CG head asks for it (the state's
 * mixed fields set as for flexible CG); convergence is accepted on a
 * recomputed residual only */
int ugk_cc_iter_mixed(const ugk_mat *K, const ugk_coarse *C, long n, const double *d_b, double *d_x, double *d_r,
                      double *d_z, double *d_p, double *d_q, const double *d_dinv, ugk_state *d_st,
                      double *d_part, void *stream);
/* Jacobi PCG on the FP32 copy of K with FP64 vectors, recurrences and dots
 * (the FP32 values are widened on load): [refresh != 0: r = b - K x in FP64
 * if a head asked for it since the last refresh -- launch it once per chunk
 * of iterations], head and p = dinv r + beta p in one kernel (a convergence
 * is only accepted on a refreshed residual), q = K32 p with (p, q), x and r
 * update with (r, dinv r).  r and the partials from ugk_fcg_start; the
 * state's mixed fields set as for flexible CG */
int ugk_mx64_iter(const ugk_mat *K, long n, const double *d_b, double *d_x, double *d_r, double *d_p, double *d_q,
                  const double *d_dinv, ugk_state *d_st, double *d_part, int refresh, void *stream);
/* z_c alone into d_zc (skipped once the solve is done); d_out NULL in
 * ugk_fcg_combine drops the network term, d_zc NULL the coarse one, and
 * ugk_fcg_head with d_x0 NULL writes no network input */
int ugk_cc_apply(const ugk_coarse *C, long n, const double *d_r, double *d_zc, const ugk_state *d_st,
                 void *stream);

/* ---------------- 3-D convolution network kernels (ugk_unet.hip) ---------- */
/* Activations are bf16 (stored as unsigned short), channels-last on a node
 * grid [nb][D][H][W][ld]: channel c of voxel v of sample b sits at
 * ((b * D + z) * H + y) * W + x) * ld + c.  Every ld and every channel
 * offset folded into a pointer is a multiple of 4. */
typedef struct { int D, H, W; } ugk_grid3;

enum { UGK_CONV_S1 = 0,   /* y[v] = sum_t W[t] x[v + off_t]                 */
       UGK_CONV_S2 = 1,   /* y[c] = sum_t W[t] x[2 c + off_t]  (coarse out)  */
       UGK_CONV_T2 = 2 }; /* y[f] = sum_t W[t] x[(f - off_t) / 2], even only */

/* Activation formats (x_f32 below): 0 bf16, 1 f32, 2 "split" -- an f32
 * tensor kept as its exact three-way bf16 split, each group of 4 channels
 * stored as hi[4], mid[4], lo[4] (24 bytes; f32 = (hi + mid) + lo exactly),
 * so the element (v, c) of a tensor of stride ld starts at bf16 offset
 * 3 (v ld + c) for c % 4 == 0. */
enum { UGK_ACT_BF16 = 0, UGK_ACT_F32 = 1, UGK_ACT_SPLIT = 2,
       UGK_ACT_F64 = 3 };  /* FP64 (the FP64 inference path only) */

/* 3x3x3 taps t = 9 (dz+1) + 3 (dy+1) + (dx+1), zero padding, no bias.
 * Activations x are bf16, f32 (fed to the bf16 matrix cores as an exact
 * three-way bf16 split made at each load) or split (the same products with
 * the split made once by the producer).  Epilogue per output element:
 * v = acc (* emul) (+ ga * gb) (f32 fields, optional), stored as bf16, or
 * f32 if y_f32. */
typedef struct ugk_conv_desc_s {
  const void *x; int ldx, cin, x_f32;      /* cin % 4 == 0, cin <= cip        */
  const unsigned short *w; int cip, cop;   /* [27][cop][cip], zero padded, %16 */
  void *y; int ldy, cout, y_f32;           /* cout <= cop                      */
  const float *emul; int ldm;              /* optional (NULL)                  */
  const float *ga; int ldga;               /* optional (NULL)                  */
  const float *gb; int ldgb;
  int mode, flip, nb;                      /* flip: use tap 26 - t             */
  ugk_grid3 gi, go;
  const int *skip;                         /* optional: no-op while *skip != 0 */
} ugk_conv_desc;
int ugk_conv(const ugk_conv_desc *d, void *stream);

/* FP64 inference of the network: the same operator (bf16 weights widened
 * exactly, f32 gates) with FP64 activations [V][C] and FP64 sums; the 16-
 * and 32-channel stages on the FP64 matrix cores.
 *   ugk_w64_frag  the matrix-core fragments of one layer (bf16 [27][cop][cip],
 *                 co outputs, ci inputs): up = 0 for the convolution, 1 for its
 *                 adjoint; ugk_n64_frag_len(nin, nout) doubles
 *   ugk_w64       wu [27][co][ci] FP64 (for ugk_n64_out)
 *   ugk_n64_down  y = (gate *) conv_s2(x): x [Vi][ldx] on gi (ci = 3, 16, 32),
 *                 y [Vo][co] on go = coarse(gi) (co = 16, 32)
 *   ugk_n64_up    y = gate h + conv_s2^T(x): x [Vc][ci] on gc (ci = 16, 32),
 *                 y, h [Vf][co] on gf with gc = coarse(gf) (co = 16, 32),
 *                 the adjoint fragments of the convolution gf -> gc
 *   ugk_n64_out   y = conv_s2^T(x) to 3 channels (ci = 16, 32), wu of that layer
 * (d_skip optional: no-op while *d_skip != 0) */
long ugk_n64_frag_len(int nin, int nout);
int ugk_w64_frag(const unsigned short *d_wf, int cop, int cip, int co, int ci, int up, double *d_frag, void *stream);
int ugk_w64(const unsigned short *d_wf, int cop, int cip, int co, int ci, double *d_wu, void *stream);
int ugk_n64_down(const double *d_x, int ldx, int ci, ugk_grid3 gi, ugk_grid3 go, const double *d_frag, int co,
                 const float *d_gate, double *d_y, const int *d_skip, void *stream);
int ugk_n64_up(const double *d_x, int ci, ugk_grid3 gc, ugk_grid3 gf, const double *d_frag, int co,
               const float *d_gate, const double *d_h, double *d_y, const int *d_skip, void *stream);
int ugk_n64_out(const double *d_x, int ci, ugk_grid3 gc, ugk_grid3 gf, const double *d_wu, double *d_y, void *stream);

/* weight gradient of y = conv(x) (mode S1 or S2), x and dy bf16 or f32
 * (rounded to bf16 for the matrix cores): block partials over ks voxel
 * slices into d_part [ks][27][cop][cip] (f32), then the fixed-order sum over
 * slices into d_grad [27][cout][cin] (added to it when accumulate) */
typedef struct {
  const void *x; int ldx, cin, x_f32;
  const void *dy; int ldd, cout, dy_f32;
  int cip, cop, mode, nb;
  ugk_grid3 gi, go;
} ugk_dw_desc;
int ugk_conv_dw(const ugk_dw_desc *d, float *d_part, int ks, float *d_grad, int accumulate, void *stream);
int ugk_conv_dw_slices(const ugk_dw_desc *d);   /* ks the call above will use */

/* f32 [n4][4] -> split format (n4 groups of 4 channels) */
int ugk_to_split(long n4, const float *d_x, unsigned short *d_y, void *stream);
/* out = a * b elementwise (f32) */
int ugk_mul(long n, const float *d_a, const float *d_b, float *d_out, const int *d_skip, void *stream);
/* gates g[v][c] = exp(b[c] - (1 + a[c]) eb[v]) for nvox voxels, C channels */
int ugk_gates(long nvox, int C, const float *d_eb, const float *d_a, const float *d_b, float *d_g,
              void *stream);
/* gate gradient: q = g h dt; d_grad_b[c] = sum q, d_grad_a[c] = -sum eb q
 * (d_part: 512 x 2C doubles; C divides 256) */
int ugk_gate_grad(long nvox, int C, const float *d_g, const float *d_h, const float *d_dt, const float *d_eb,
                  double *d_part, float *d_grad_a, float *d_grad_b, void *stream);
/* energy loss of one training pair (r, e) scaled by s, network output y
 * [n/3][4]: z = dinv rr + mask sqrt(dinv) y; then, after kz = K z,
 * loss partials sum (ee - z)(rr - kz) and dy = coef mask sqrt(dinv) (rr - kz);
 * returns the number of partials */
int ugk_el_z(long n, const float *d_rr, const float *d_dinv, const float *d_mask, const float *d_y,
             float *d_z, void *stream);
int ugk_el_grad(long n, const float *d_rr, const float *d_ee, const float *d_z, const float *d_kz,
                const float *d_dinv, const float *d_mask, float coef, float *d_dy, double *d_lpart,
                void *stream);
/* Adam on one f32 weight tensor [27][cout][cin] (or a vector: 1 x P x 1);
 * refreshes the bf16 forward [27][cop][cip] and transposed [27][cip][cop]
 * copies when d_wf is not NULL.  lr <= 0: copies only */
int ugk_adam(long P, int cout, int cin, int cop, int cip, float *d_w, const float *d_g, float *d_m,
             float *d_v, unsigned short *d_wf, unsigned short *d_wt, float lr, float b1, float b2,
             float eps, long step, void *stream);
/* training pair from a PCG snapshot (x, r) and the converged u: per-block
 * partials of (sum dinv r^2, sum (u - x) r) (returns the blocks); then, with
 * s = rms(sqrt(dinv) r): V[v] = mask sqrt(dinv) r / s, rr = r / s,
 * ee = (u - x) / s (f32) */
int ugk_pair_stats(long n, const double *d_x, const double *d_r, const double *d_u,
                   const double *d_dinv, double *d_part, void *stream);
int ugk_pair_write(long n, const double *d_x, const double *d_r, const double *d_u,
                   const double *d_dinv, const float *d_mask, double s, float *d_V, float *d_rr,
                   float *d_ee, void *stream);

#ifdef __cplusplus
}
#endif
#endif
