// ugk_kernels.hip -- HIP kernels for gfx950 behind the thin C-ABI in ugk.h.
//
// Hot path of uFEM's linear-static solve on one MI355X:
//   * element 9 (8-node brick) stiffness, one wavefront per element, Gauss
//     gradients staged in LDS, K_ab = sum_gp w detJ (lam g_a g_b^T +
//     mu g_b g_a^T + mu (g_a.g_b) I) per node pair -- the isotropic form of
//     H^T D H of fem/fem_e009.c:587-602 with D of fem/fem_pl3d.c:32-58;
//   * deterministic assembly: each colour's elements share no node, so every
//     3x3 block of K is updated by exactly one lane per launch (no atomics),
//     and colours run in a fixed order;
//   * SELL-64 block SpMV streamed from HBM with fully coalesced value loads;
//   * Jacobi PCG with the p-update folded into the SpMV gather and the vector
//     update fused with its three dot products; reductions finish in the last
//     block to arrive (fixed summation order, so results repeat bit for bit).
#include <hip/hip_runtime.h>
#include <hip/hip_cooperative_groups.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "ugk.h"

namespace cg = cooperative_groups;

static thread_local char g_err[512] = "";

static int fail(const char *what, hipError_t e)
{
  snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
  return -1;
}
// error reporting for the other kernel sources of the library
int ugk_hip_fail(const char *what, hipError_t e) { return fail(what, e); }
#define CK(call)                                  \
  do {                                            \
    hipError_t e_ = (call);                       \
    if (e_ != hipSuccess) return fail(#call, e_); \
  } while (0)
#define CK_LAUNCH(what) CK(hipGetLastError())

#define WG 256
#define NW (WG / 64)

// --------------------------------------------------------------------------
// reductions
// --------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}

__device__ __forceinline__ double ld_agent(const double *p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(double *p, double v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum NV values over the block (fixed order), publish the block's partials,
// and let the last block to arrive reduce all partials in index order.
// Returns true in every thread of the last block, with tot[] valid in all
// its threads.  MAXR: reduce with max instead of sum.
template <int NV, bool MAXR = false>
__device__ bool block_reduce_last(double (&v)[NV], double *part, unsigned int *counter,
                                  double (&tot)[NV])
{
  __shared__ double sh[NV][NW];
  __shared__ int is_last;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned G = gridDim.x;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = MAXR ? wave_max(v[k]) : wave_sum(v[k]);
    if (lane == 0) sh[k][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double a = sh[k][0];
      for (int w = 1; w < NW; w++) a = MAXR ? fmax(a, sh[k][w]) : a + sh[k][w];
      st_agent(&part[(size_t)k * G + blockIdx.x], a);
    }
    __threadfence();
    unsigned t = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    is_last = (t == G - 1);
  }
  __syncthreads();
  if (!is_last) return false;
  __threadfence();
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double a = MAXR ? 0.0 : 0.0;
    for (unsigned i = threadIdx.x; i < G; i += WG) {
      double pv = ld_agent(&part[(size_t)k * G + i]);
      a = MAXR ? fmax(a, pv) : a + pv;
    }
    double s = MAXR ? wave_max(a) : wave_sum(a);
    __syncthreads();
    if (lane == 0) sh[k][wave] = s;
    __syncthreads();
    double r = sh[k][0];
    for (int w = 1; w < NW; w++) r = MAXR ? fmax(r, sh[k][w]) : r + sh[k][w];
    tot[k] = r;
  }
  if (threadIdx.x == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

// Redundant grid reduction inside a cooperative kernel: every block sums the
// G published partials in index order (identical result in every block).
template <int NV>
__device__ void grid_partials_sum(const double *part, unsigned G, double (&tot)[NV])
{
  __shared__ double sh[NV][NW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double a = 0.0;
    for (unsigned i = threadIdx.x; i < G; i += WG) a += part[(size_t)k * G + i];  // after grid.sync
    double s = wave_sum(a);
    if (lane == 0) sh[k][wave] = s;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double r = sh[k][0];
    for (int w = 1; w < NW; w++) r += sh[k][w];
    tot[k] = r;
  }
  __syncthreads();
}

template <int NV>
__device__ void block_publish(double (&v)[NV], double *part)
{
  __shared__ double sh[NV][NW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = wave_sum(v[k]);
    if (lane == 0) sh[k][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double a = sh[k][0];
      for (int w = 1; w < NW; w++) a += sh[k][w];
      part[(size_t)k * gridDim.x + blockIdx.x] = a;   // published by the next grid.sync
    }
  }
  __syncthreads();
}

// Block partials with plain stores.  The consumer is always a later kernel
// on the same stream, whose start acquires what this kernel released at its
// end, so no fences or atomics are needed on the iteration's hot path.
template <int NV>
__device__ void block_partials(double (&v)[NV], double *part)
{
  __shared__ double sh[NV][NW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = wave_sum(v[k]);
    if (lane == 0) sh[k][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double a = sh[threadIdx.x][0];
    for (int w = 1; w < NW; w++) a += sh[threadIdx.x][w];
    part[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = a;
  }
}

// A thread's share of the partials of NV values (grid G, stride TB), with
// RP_U loads of every value in flight per pass (a loop that waits on each
// load costs one memory round trip per entry: the partials were written by
// other XCDs); fixed order, so every block gets the same sums.
template <int NV, int TB>
__device__ __forceinline__ void thread_partials(const double *__restrict__ part, unsigned G, double (&a)[NV])
{
  constexpr int RP_U = NV >= 4 ? 4 : 8;
#pragma unroll
  for (int k = 0; k < NV; k++) a[k] = 0.0;
  unsigned i = threadIdx.x;
  for (; i + (RP_U - 1) * TB < G; i += RP_U * TB) {
    double t[NV][RP_U];
#pragma unroll
    for (int k = 0; k < NV; k++)
#pragma unroll
      for (int u = 0; u < RP_U; u++) t[k][u] = part[(size_t)k * G + i + u * TB];
#pragma unroll
    for (int k = 0; k < NV; k++)
#pragma unroll
      for (int u = 0; u < RP_U; u++) a[k] += t[k][u];
  }
  double t[NV][RP_U];
#pragma unroll
  for (int u = 0; u < RP_U; u++)
#pragma unroll
    for (int k = 0; k < NV; k++) {
      const bool in = i + u * TB < G;        // the address stays in the region either way
      const double pv = part[(size_t)k * G + (in ? i + u * TB : 0)];
      t[k][u] = in ? pv : 0.0;
    }
#pragma unroll
  for (int k = 0; k < NV; k++)
#pragma unroll
    for (int u = 0; u < RP_U; u++) a[k] += t[k][u];
}

// Every block sums the partials a previous kernel (grid G) published, in
// a fixed order, so all blocks hold bit-identical totals.
template <int NV>
__device__ void reduce_prev(const double *__restrict__ part, unsigned G, double (&tot)[NV])
{
  __shared__ double sh[NV][NW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double a[NV];
  thread_partials<NV, WG>(part, G, a);
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = wave_sum(a[k]);
    if (lane == 0) sh[k][wave] = s;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double r = sh[k][0];
    for (int w = 1; w < NW; w++) r += sh[k][w];
    tot[k] = r;
  }
}

// reduce_prev and block_partials for blocks of TB threads
template <int NV, int TB>
__device__ void reduce_prev_tb(const double *__restrict__ part, unsigned G, double (&tot)[NV])
{
  __shared__ double sh[NV][TB / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double a[NV];
  thread_partials<NV, TB>(part, G, a);
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = wave_sum(a[k]);
    if (lane == 0) sh[k][wave] = s;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double r = sh[k][0];
    for (int w = 1; w < TB / 64; w++) r += sh[k][w];
    tot[k] = r;
  }
}
template <int NV, int TB>
__device__ void block_partials_tb(double (&v)[NV], double *part)
{
  __shared__ double sh[NV][TB / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = wave_sum(v[k]);
    if (lane == 0) sh[k][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double a = sh[threadIdx.x][0];
    for (int w = 1; w < TB / 64; w++) a += sh[threadIdx.x][w];
    part[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = a;
  }
}

// convergence test shared by the PCG kernels
__device__ __forceinline__ int pcg_test(const ugk_state *st, long k, double rr, double xx, double bb,
                                        double rz)
{
  const double nr = sqrt(rr), nb = sqrt(bb);
  if (!(nb > 0.0)) return 1;
  if (!(rz == rz) || !(rr == rr)) return 2;
  if (st->rule == 1 ? nr <= st->tol * nb : nr <= st->tol * (st->normA * sqrt(xx) + nb)) return 1;
  if (k >= st->maxit) return 3;
  return 0;
}

// --------------------------------------------------------------------------
// SELL-64 addressing
// --------------------------------------------------------------------------
struct SellView {
  const int *slen;
  const long *soff;
  const int *rowlen;
  const int *cols;
  long n_brows, n_slices;
  const float *corr;  // FP32 copy's diagonal correction [9][n_brows], or null
  const unsigned short *rem;  // FP32 copy's bf16 remainders (same positions), or null
};

__device__ __forceinline__ int find_slot(const SellView &M, long I, long J)
{
  // binary search of column J among the sorted real slots of block row I
  const long sl = I >> 6, ln = I & 63;
  const int *c = M.cols + M.soff[sl] * 64 + ln;
  int lo = 0, hi = M.rowlen[I] - 1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    int cm = c[(long)mid * 64];
    if (cm == J) return mid;
    if (cm < J) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

// Position of value q (= BS r + c, row-major in the block) of lane l in slot
// S.  A 3x3 slot holds 576 values: each lane's first eight values in adjacent
// pairs (pair k of lane l at 128 k + 2 l, one 16-byte load in FP64), the
// ninth in a 64-wide tail at 512 + l.  A 1x1 slot holds one value per lane.
// The host-side twin is ugk_vpos3 in ugk.h.
template <int BS>
__device__ __forceinline__ long vpos(long S, int q, int l)
{
  if (BS == 1) return S * 64 + l;
  return S * 576 + (q < 8 ? (q >> 1) * 128 + 2 * l + (q & 1) : 512 + l);
}

template <typename T>
__device__ __forceinline__ T ldnt(const T *p)
{
  return __builtin_nontemporal_load(p);
}

template <typename T> struct Pair;
template <> struct Pair<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct Pair<float> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct Pair<unsigned short> { typedef unsigned short type __attribute__((ext_vector_type(2))); };

template <typename T> struct Quad;
template <> struct Quad<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct Quad<unsigned short> { typedef unsigned short type __attribute__((ext_vector_type(4))); };

// Position of value q (row-major in the 3x3 block) of lane l in slot S of the
// FP32 copy and of its bf16 remainders: the first eight in two runs of four
// (run k of lane l at 256 k + 4 l: one 16-byte load in FP32), the ninth in
// the 64-wide tail at 512 + l.  (The FP64 values keep pairs, ugk_vpos3.)
__device__ __forceinline__ long vpos32(long S, int q, int l)
{
  return S * 576 + (q < 8 ? (q >> 2) * 256 + 4 * l + (q & 3) : 512 + l);
}

// the BS*BS values of lane `lane` in slot S, row-major (NT: non-temporal loads)
template <int BS, bool NT, typename T>
__device__ __forceinline__ void ld_block(const T *__restrict__ vals, long S, int lane, T (&b)[BS * BS])
{
  if constexpr (BS == 1) {
    b[0] = NT ? ldnt(vals + S * 64 + lane) : vals[S * 64 + lane];
  } else if constexpr (sizeof(T) <= 4) {   // the FP32 copy's layout (vpos32)
    typedef typename Quad<T>::type T4;
    const T *p = vals + S * 576;
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const T4 *a = (const T4 *)(p + k * 256 + 4 * lane);
      const T4 t = NT ? ldnt(a) : *a;
      b[4 * k] = t.x;
      b[4 * k + 1] = t.y;
      b[4 * k + 2] = t.z;
      b[4 * k + 3] = t.w;
    }
    b[8] = NT ? ldnt(p + 512 + lane) : p[512 + lane];
  } else {
    typedef typename Pair<T>::type T2;
    const T *p = vals + S * 576;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const T2 *a = (const T2 *)(p + k * 128 + 2 * lane);
      const T2 t = NT ? ldnt(a) : *a;
      b[2 * k] = t.x;
      b[2 * k + 1] = t.y;
    }
    b[8] = NT ? ldnt(p + 512 + lane) : p[512 + lane];
  }
}

// logical slice for a block: XCD-aware remap so that each XCD (blocks are
// dealt round-robin over 8 XCDs) streams one contiguous range of rows and
// its L2 keeps the x entries those rows gather
__device__ __forceinline__ long remap_block(long b, long nb)
{
  // nb is a multiple of 8 (see spmv_blocks), so this is a bijection
  const long per = nb >> 3;
  return (b & 7) * per + (b >> 3);
}

// This block's contiguous range [i1, i2) of [0, n) for vector kernels: block
// chunks in remapped order, so the chunks of one XCD form one range, as the
// rows of the SpMV kernels do, and the entries an XCD's SpMV gathers or
// wrote stay in its L2 between the kernels of an iteration
__device__ __forceinline__ void xcd_chunk(long n, long &i1, long &i2)
{
  const long lb = (gridDim.x & 7) ? (long)blockIdx.x : remap_block(blockIdx.x, gridDim.x);
  // a 32-bit division where n allows (a 64-bit one is a long software sequence)
  const long chunk = n < (1L << 32) - gridDim.x ? (long)(((unsigned)n + gridDim.x - 1) / gridDim.x)
                                                : (n + gridDim.x - 1) / gridDim.x;
  i1 = lb * chunk;
  i2 = min(n, i1 + chunk);
}

// The same with an even chunk, for kernels that move two entries per lane
// with 16-byte accesses (chunk starts stay 16-byte aligned)
__device__ __forceinline__ void xcd_chunk_even(long n, long &i1, long &i2)
{
  const long lb = (gridDim.x & 7) ? (long)blockIdx.x : remap_block(blockIdx.x, gridDim.x);
  const long h = (n + 1) / 2;
  const long c2 = h < (1L << 32) - gridDim.x ? (long)(((unsigned)h + gridDim.x - 1) / gridDim.x)
                                             : (h + gridDim.x - 1) / gridDim.x;
  i1 = min(n, lb * 2 * c2);
  i2 = min(n, i1 + 2 * c2);
}
typedef double dbl2 __attribute__((ext_vector_type(2)));

// --------------------------------------------------------------------------
// element stiffness
// --------------------------------------------------------------------------
template <int NPTS>
struct Gauss;
template <>
struct Gauss<2> {
  __device__ static double pt(int i) { return i == 0 ? -0.57735026918962576451 : 0.57735026918962576451; }
  __device__ static double wt(int) { return 1.0; }
};
template <>
struct Gauss<3> {
  __device__ static double pt(int i) { return i == 0 ? -0.77459666924148337704 : (i == 1 ? 0.0 : 0.77459666924148337704); }
  __device__ static double wt(int i) { return i == 1 ? 0.88888888888888888889 : 0.55555555555555555556; }
};

// corner signs of the reference brick, fem/fem_e009.c:36-69
__device__ __forceinline__ void corner(int n, double &xs, double &ys, double &zs)
{
  xs = ((n + 1) & 2) ? 1.0 : -1.0;   // 0:-,1:+,2:+,3:-
  ys = (n & 2) ? 1.0 : -1.0;         // 0:-,1:-,2:+,3:+
  zs = (n & 4) ? 1.0 : -1.0;
}

// physical shape-function gradients g[gp][n][a] = dN_n/dx_a and weights
// w[gp] = detJ w_i w_j w_k of one element (xyz: 8 corners x 3, all in LDS),
// spread over the 64 lanes of its wave; detJ <= 0 flags st->err = 1
template <int NPTS>
__device__ __forceinline__ void elem_gradients(const double *xyz, double *g, double *w, int lane, ugk_state *st)
{
  constexpr int NGP = NPTS * NPTS * NPTS;
  for (int t = lane; t < NGP * 8; t += 64) {
    const int gp = t >> 3, n = t & 7;
    const int gi = gp / (NPTS * NPTS), gj = (gp / NPTS) % NPTS, gk = gp % NPTS;
    const double xi = Gauss<NPTS>::pt(gi), eta = Gauss<NPTS>::pt(gj), zeta = Gauss<NPTS>::pt(gk);
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double dn[3];
#pragma unroll
    for (int m = 0; m < 8; m++) {
      double xs, ys, zs;
      corner(m, xs, ys, zs);
      const double d0 = 0.125 * xs * (1.0 + ys * eta) * (1.0 + zs * zeta);
      const double d1 = 0.125 * ys * (1.0 + xs * xi) * (1.0 + zs * zeta);
      const double d2 = 0.125 * zs * (1.0 + xs * xi) * (1.0 + ys * eta);
      if (m == n) { dn[0] = d0; dn[1] = d1; dn[2] = d2; }
#pragma unroll
      for (int b = 0; b < 3; b++) {
        const double xb = xyz[3 * m + b];
        J[0][b] += d0 * xb;
        J[1][b] += d1 * xb;
        J[2][b] += d2 * xb;
      }
    }
    // J[a][b] = dx_b / dxi_a ; inverse by adjugate
    const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    if (!(det > 0.0)) st->err = 1;
    const double id = 1.0 / det;
    double Ji[3][3];
    Ji[0][0] = c00 * id;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    Ji[1][0] = c01 * id;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    Ji[2][0] = c02 * id;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
#pragma unroll
    for (int a = 0; a < 3; a++) g[(gp * 8 + n) * 3 + a] = Ji[a][0] * dn[0] + Ji[a][1] * dn[1] + Ji[a][2] * dn[2];
    if (n == 0) w[gp] = det * Gauss<NPTS>::wt(gi) * Gauss<NPTS>::wt(gj) * Gauss<NPTS>::wt(gk);
  }
}

template <int NPTS, bool DENSE>
__global__ __launch_bounds__(WG) void k_elem_stiff(long n_list, const int *__restrict__ elist,
                                                   const double *__restrict__ cx,
                                                   const double *__restrict__ cy,
                                                   const double *__restrict__ cz,
                                                   const int *__restrict__ enodes,
                                                   const double *__restrict__ Ev,
                                                   const double *__restrict__ nuv,
                                                   double *__restrict__ Ke, SellView M,
                                                   double *__restrict__ vals, ugk_state *st)
{
  constexpr int NGP = NPTS * NPTS * NPTS;
  __shared__ double s_xyz[NW][24];
  __shared__ double s_g[NW][NGP][8][3];
  __shared__ double s_w[NW][NGP];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long idx = (long)blockIdx.x * NW + wave;
  const bool active = idx < n_list;
  const long e = active ? (DENSE ? idx : (long)elist[idx]) : 0;

  if (active && lane < 24) {
    const int nd = lane / 3, c = lane % 3;
    if (DENSE) {
      s_xyz[wave][lane] = cx[e * 24 + lane];
    } else {
      const int node = enodes[e * 8 + nd];
      s_xyz[wave][lane] = c == 0 ? cx[node] : (c == 1 ? cy[node] : cz[node]);
    }
  }
  __syncthreads();
  if (active) elem_gradients<NPTS>(s_xyz[wave], &s_g[wave][0][0][0], s_w[wave], lane, st);
  __syncthreads();
  if (!active) return;
  const int a = lane >> 3, b = lane & 7;
  const double E = Ev[e], nu = nuv[e];
  const double lam = (E * nu) / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double mu = E / (2.0 * (1.0 + nu));
  double k[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int gp = 0; gp < NGP; gp++) {
    const double w = s_w[wave][gp];
    const double ga0 = s_g[wave][gp][a][0], ga1 = s_g[wave][gp][a][1], ga2 = s_g[wave][gp][a][2];
    const double gb0 = s_g[wave][gp][b][0], gb1 = s_g[wave][gp][b][1], gb2 = s_g[wave][gp][b][2];
    const double wl = w * lam, wm = w * mu;
    const double dd = wm * (ga0 * gb0 + ga1 * gb1 + ga2 * gb2);
    k[0] += wl * ga0 * gb0 + wm * ga0 * gb0 + dd;
    k[1] += wl * ga0 * gb1 + wm * ga1 * gb0;
    k[2] += wl * ga0 * gb2 + wm * ga2 * gb0;
    k[3] += wl * ga1 * gb0 + wm * ga0 * gb1;
    k[4] += wl * ga1 * gb1 + wm * ga1 * gb1 + dd;
    k[5] += wl * ga1 * gb2 + wm * ga2 * gb1;
    k[6] += wl * ga2 * gb0 + wm * ga0 * gb2;
    k[7] += wl * ga2 * gb1 + wm * ga1 * gb2;
    k[8] += wl * ga2 * gb2 + wm * ga2 * gb2 + dd;
  }
  if (DENSE) {
    double *o = Ke + e * 576;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) o[(3 * a + i) * 24 + 3 * b + j] = k[3 * i + j];
  } else {
    const long I = enodes[e * 8 + a], J = enodes[e * 8 + b];
    const int s = find_slot(M, I, J);
    if (s < 0) { st->err = 3; return; }
    const long S = M.soff[I >> 6] + s;
#pragma unroll
    for (int q = 0; q < 9; q++) vals[vpos<3>(S, q, (int)(I & 63))] += k[q];
  }
}

// strains and stresses at the Gauss points of every element (the results
// pass of e009_stiff Mode=AF_YES for a linear material, fem/fem_e009.c:417-535):
// epsilon = H u_e in the order x, y, z, yz, zx, xy (engineering shears),
// sigma = D epsilon with the isotropic D of fem/fem_pl3d.c:32-58; out
// [e][gp][12] = sigma (6) then epsilon (6), the order of the element's
// repeating results (fem/fem_e009.c:881-883); gp = (i NPTS + j) NPTS + k
template <int NPTS>
__global__ __launch_bounds__(WG) void k_elem_results(long ne, const double *__restrict__ cx,
                                                     const double *__restrict__ cy,
                                                     const double *__restrict__ cz,
                                                     const int *__restrict__ enodes,
                                                     const double *__restrict__ Ev,
                                                     const double *__restrict__ nuv,
                                                     const double *__restrict__ u, double *__restrict__ out,
                                                     ugk_state *st)
{
  constexpr int NGP = NPTS * NPTS * NPTS;
  __shared__ double s_xyz[NW][24];
  __shared__ double s_u[NW][24];
  __shared__ double s_g[NW][NGP][8][3];
  __shared__ double s_w[NW][NGP];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long e = (long)blockIdx.x * NW + wave;
  const bool active = e < ne;
  if (active && lane < 24) {
    const int node = enodes[e * 8 + lane / 3], c = lane % 3;
    s_xyz[wave][lane] = c == 0 ? cx[node] : (c == 1 ? cy[node] : cz[node]);
    s_u[wave][lane] = u[(long)node * 3 + c];
  }
  __syncthreads();
  if (active) elem_gradients<NPTS>(s_xyz[wave], &s_g[wave][0][0][0], s_w[wave], lane, st);
  __syncthreads();
  if (!active) return;
  const double E = Ev[e], nu = nuv[e];
  const double d1 = E * (1.0 - nu) / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double d2 = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double d3 = E / (2.0 * (1.0 + nu));
  for (int gp = lane; gp < NGP; gp += 64) {
    double ex = 0, ey = 0, ez = 0, gyz = 0, gzx = 0, gxy = 0;
#pragma unroll
    for (int n = 0; n < 8; n++) {
      const double g0 = s_g[wave][gp][n][0], g1 = s_g[wave][gp][n][1], g2 = s_g[wave][gp][n][2];
      const double u0 = s_u[wave][3 * n], u1 = s_u[wave][3 * n + 1], u2 = s_u[wave][3 * n + 2];
      ex += g0 * u0;
      ey += g1 * u1;
      ez += g2 * u2;
      gyz += g2 * u1 + g1 * u2;
      gzx += g0 * u2 + g2 * u0;
      gxy += g1 * u0 + g0 * u1;
    }
    double *o = out + (e * NGP + gp) * 12;
    o[0] = d1 * ex + d2 * (ey + ez);
    o[1] = d1 * ey + d2 * (ex + ez);
    o[2] = d1 * ez + d2 * (ex + ey);
    o[3] = d3 * gyz;
    o[4] = d3 * gzx;
    o[5] = d3 * gxy;
    o[6] = ex;
    o[7] = ey;
    o[8] = ez;
    o[9] = gyz;
    o[10] = gzx;
    o[11] = gxy;
  }
}

// --------------------------------------------------------------------------
// supports and reactions (block size 3)
// --------------------------------------------------------------------------
__global__ void k_save_rows(SellView M, const double *__restrict__ vals, long n_fixed,
                            const int *__restrict__ fixed, int maxb, double *rows, int *rcols)
{
  long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n_fixed) return;
  const long d = fixed[f], I = d / 3, r = d % 3;
  const long so = M.soff[I >> 6], ln = I & 63;
  const int L = M.rowlen[I];
  for (int s = 0; s < maxb; s++) {
    if (s < L) {
      rcols[f * maxb + s] = M.cols[(so + s) * 64 + ln];
      for (int c = 0; c < 3; c++) rows[(f * maxb + s) * 3 + c] = vals[vpos<3>(so + s, (int)(3 * r + c), (int)ln)];
    } else {
      rcols[f * maxb + s] = -1;
      for (int c = 0; c < 3; c++) rows[(f * maxb + s) * 3 + c] = 0.0;
    }
  }
}

__global__ void k_zero_rowcol(SellView M, double *__restrict__ vals, long n_fixed,
                              const int *__restrict__ fixed)
{
  long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n_fixed) return;
  const long d = fixed[f], I = d / 3, r = d % 3;
  const long so = M.soff[I >> 6], ln = I & 63;
  const int L = M.rowlen[I];
  for (int s = 0; s < L; s++) {
    for (int c = 0; c < 3; c++) vals[vpos<3>(so + s, (int)(3 * r + c), (int)ln)] = 0.0;  // row d
    const long Jn = M.cols[(so + s) * 64 + ln];                                     // column d
    const int t = find_slot(M, Jn, I);
    if (t < 0) continue;
    const long soJ = M.soff[Jn >> 6], lnJ = Jn & 63;
    for (int rr = 0; rr < 3; rr++) vals[vpos<3>(soJ + t, (int)(3 * rr + r), (int)lnJ)] = 0.0;
  }
}

__global__ void k_set_diag_one(SellView M, double *__restrict__ vals, long n_fixed,
                               const int *__restrict__ fixed)
{
  long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n_fixed) return;
  const long d = fixed[f], I = d / 3, r = d % 3;
  const int s = find_slot(M, I, I);
  if (s < 0) return;
  vals[vpos<3>(M.soff[I >> 6] + s, (int)(4 * r), (int)(I & 63))] = 1.0;
}

__global__ void k_reactions(long n_fixed, const int *__restrict__ fixed, int maxb,
                            const double *__restrict__ rows, const int *__restrict__ rcols,
                            const double *__restrict__ u, const double *__restrict__ Forig,
                            double *R)
{
  long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n_fixed) return;
  double acc = 0.0;
  for (int s = 0; s < maxb; s++) {
    const int J = rcols[f * maxb + s];
    if (J < 0) break;
    for (int c = 0; c < 3; c++) acc += rows[(f * maxb + s) * 3 + c] * u[3L * J + c];
  }
  R[f] = acc - (Forig ? Forig[fixed[f]] : 0.0);
}

// --------------------------------------------------------------------------
// SpMV and matrix utilities
// --------------------------------------------------------------------------
template <int BS, typename T, typename TX>
__device__ __forceinline__ void sell_row(const SellView &M, const T *__restrict__ vals, long sl,
                                         int lane, const TX *__restrict__ x, TX (&acc)[BS])
{
  const int L = M.slen[sl];
  const long so = M.soff[sl];
  const int *cp = M.cols + so * 64 + lane;
#pragma unroll
  for (int r = 0; r < BS; r++) acc[r] = 0;
#pragma unroll 3
  for (int s = 0; s < L; s++) {
    const long J = cp[(long)s * 64];
    TX xv[BS];
#pragma unroll
    for (int c = 0; c < BS; c++) xv[c] = x[J * BS + c];
    T v[BS * BS];
    ld_block<BS, false>(vals, so + s, lane, v);
#pragma unroll
    for (int r = 0; r < BS; r++)
#pragma unroll
      for (int c = 0; c < BS; c++) acc[r] += (TX)v[r * BS + c] * xv[c];
  }
}

// Slot-split SELL row: a workgroup of 4 waves serves 2 slices; the two waves
// of a slice stream the first and second half of its slots (more loads in
// flight per row), the second half's sums meet the first's in LDS.  Matrix
// values and columns are read once, so they are loaded non-temporally to
// leave L2 to the gathered vector.  Returns true in the lanes that hold a
// finished row sum (first-half wave, real row); `row` is the block row.
template <int BS, typename T, typename TX>
__device__ __forceinline__ bool sell_row_split(const SellView &M, const T *__restrict__ vals,
                                               const TX *__restrict__ x, TX (&acc)[BS], long &row)
{
  __shared__ TX sh[2][BS][64];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, half = w & 1, pair = w >> 1;
  const long lb = remap_block(blockIdx.x, gridDim.x);
  const long sl = lb * 2 + pair;
#pragma unroll
  for (int r = 0; r < BS; r++) acc[r] = 0;
  if (sl < M.n_slices) {
    const int L = M.slen[sl];
    const int h = (L + 1) >> 1;
    const int s0 = half ? h : 0, s1 = half ? L : h;
    const long so = M.soff[sl];
    const int *cp = M.cols + so * 64 + lane;
    for (int s = s0; s < s1; s++) {
      const long J = ldnt(cp + (long)s * 64);
      TX xv[BS];
#pragma unroll
      for (int c = 0; c < BS; c++) xv[c] = x[J * BS + c];
      T v[BS * BS];
      ld_block<BS, true>(vals, so + s, lane, v);
      if constexpr (sizeof(T) == 4 && sizeof(TX) == 8) {
        if (M.rem) {   // 48-bit values: FP32 + bf16 remainder, summed exactly in FP64
          unsigned short u[BS * BS];
          ld_block<BS, true>(M.rem, so + s, lane, u);
#pragma unroll
          for (int r = 0; r < BS; r++)
#pragma unroll
            for (int c = 0; c < BS; c++)
              acc[r] += ((TX)v[r * BS + c] + (TX)__uint_as_float((unsigned)u[r * BS + c] << 16)) * xv[c];
          continue;
        }
      }
#pragma unroll
      for (int r = 0; r < BS; r++)
#pragma unroll
        for (int c = 0; c < BS; c++) acc[r] += (TX)v[r * BS + c] * xv[c];
    }
  }
  row = sl * 64 + lane;
  if (half) {
    if constexpr (BS == 3 && sizeof(T) == 4 && sizeof(TX) == 8) {
      // FP32 copy with FP64 vectors: add C_I x_I (C_I: the block row's summed
      // rounding errors, symmetrised), which keeps rigid translations, where
      // a stiffness matrix's rows cancel, cancelled (with FP32 vectors their
      // own rounding dominates, and the term is left out).  The second-half
      // wave has no more slots than the first, so it takes this term.
      if (M.corr && sl < M.n_slices && row < M.n_brows) {
        TX xi[3];
#pragma unroll
        for (int c = 0; c < 3; c++) xi[c] = x[row * 3 + c];
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
          for (int c = 0; c < 3; c++) acc[r] += (TX)M.corr[(r * 3 + c) * M.n_brows + row] * xi[c];
      }
    }
#pragma unroll
    for (int r = 0; r < BS; r++) sh[pair][r][lane] = acc[r];
  }
  __syncthreads();
  if (half || sl >= M.n_slices) return false;
#pragma unroll
  for (int r = 0; r < BS; r++) acc[r] += sh[pair][r][lane];
  return row < M.n_brows;
}

template <int BS, typename T, typename TX>
__global__ __launch_bounds__(WG) void k_spmv(SellView M, const T *__restrict__ vals,
                                             const TX *__restrict__ x, TX *__restrict__ y)
{
  TX acc[BS];
  long row;
  if (sell_row_split<BS, T, TX>(M, vals, x, acc, row)) {
#pragma unroll
    for (int r = 0; r < BS; r++) y[row * BS + r] = acc[r];
  }
}

template <int BS>
__global__ void k_diag_inv(SellView M, const double *__restrict__ vals, double *dinv, ugk_state *st)
{
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M.n_brows * BS) return;
  const long I = i / BS, r = i % BS;
  const int s = find_slot(M, I, I);
  double v = 0.0;
  if (s >= 0) v = vals[vpos<BS>(M.soff[I >> 6] + s, (int)(r * BS + r), (int)(I & 63))];
  if (!(fabs(v) >= 1e-17)) { st->err = 2; dinv[i] = 0.0; return; }
  dinv[i] = 1.0 / v;
}

template <int BS>
__global__ __launch_bounds__(WG) void k_mat_norm(SellView M, const double *__restrict__ vals,
                                                 ugk_state *st, double *part)
{
  double v[1] = {0.0}, tot[1];
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < M.n_brows * BS; i += (long)gridDim.x * WG) {
    const long I = i / BS, r = i % BS;
    const long so = M.soff[I >> 6], ln = I & 63;
    double s2 = 0.0;
    for (int s = 0; s < M.rowlen[I]; s++)
      for (int c = 0; c < BS; c++) {
        const double a = vals[vpos<BS>(so + s, (int)(r * BS + c), (int)ln)];
        s2 += a * a;
      }
    v[0] = fmax(v[0], sqrt(s2));
  }
  if (block_reduce_last<1, true>(v, part, &st->counter[7], tot) && threadIdx.x == 0) st->normA = tot[0];
}

template <typename T>
__global__ void k_vals_f32(long n, const double *__restrict__ a, T *__restrict__ b)
{
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    b[i] = (T)a[i];
}

// FP64 value index (pairs layout of a 3x3-block slot) -> its FP32 copy index (vpos32)
__device__ __forceinline__ long vidx64_to32(long i)
{
  const long S = i / 576;
  const int off = (int)(i - S * 576);
  if (off >= 512) return i;
  const int k = off >> 7, l = (off & 127) >> 1, q = 2 * k + (off & 1);
  return vpos32(S, q, l);
}

// FP32 copy of 3x3-block values, in the copy's own layout
__global__ void k_vals_f32_b3(long n, const double *__restrict__ a, float *__restrict__ b)
{
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    b[vidx64_to32(i)] = (float)a[i];
}

// bf16 (round to nearest even) of v - fl32(v): with it the FP32 copy holds
// 48-bit values (about 2e-10 relative); 3x3 blocks (B3) in the copy's layout
template <bool B3>
__global__ void k_vals_rem16(long n, const double *__restrict__ a, unsigned short *__restrict__ b)
{
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float r = (float)(a[i] - (double)(float)a[i]);
    unsigned u = __float_as_uint(r);
    u += 0x7fffu + ((u >> 16) & 1u);
    b[B3 ? vidx64_to32(i) : i] = (unsigned short)(u >> 16);
  }
}

// Diagonal correction of the FP32 copy of a 3x3-block matrix: for block row I,
// C = sum over its blocks J of (K_IJ - fl32(K_IJ)) (minus the bf16 remainder
// rem, when the copy keeps one), kept as (C + C^T) / 2 so
// the corrected matrix stays symmetric.  A stiffness matrix's block rows sum
// to (nearly) zero: fl32 alone breaks that cancellation by about 1e-7 |K| |x|
// on smooth fields, which stalls the true residual near 1e-3; with C the error
// on a smooth field falls with its differences x_J - x_I instead.
__global__ void k_corr32(SellView M, const double *__restrict__ vals, const unsigned short *__restrict__ rem,
                         float *__restrict__ corr)
{
  const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= M.n_brows) return;
  const long so = M.soff[I >> 6];
  const int ln = (int)(I & 63), L = M.rowlen[I];
  double c[9];
#pragma unroll
  for (int q = 0; q < 9; q++) c[q] = 0.0;
  for (int s = 0; s < L; s++) {
#pragma unroll
    for (int q = 0; q < 9; q++) {
      const double v = vals[vpos<3>(so + s, q, ln)], hi = (double)(float)v;
      c[q] += v - (rem ? hi + (double)__uint_as_float((unsigned)rem[vpos32(so + s, q, ln)] << 16) : hi);
    }
  }
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int q = 0; q < 3; q++) corr[(r * 3 + q) * M.n_brows + I] = (float)(0.5 * (c[r * 3 + q] + c[q * 3 + r]));
}

// dst[map[i]] = src[i]: host-ordered values into their device slots
__global__ void k_scatter_vals(long n, const int *__restrict__ map, const double *__restrict__ src,
                               double *__restrict__ dst)
{
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    dst[map[i]] = src[i];
}

__global__ void k_fill_zero(long nwords, unsigned int *p)
{
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += (long)gridDim.x * blockDim.x)
    p[i] = 0u;
}

__global__ __launch_bounds__(WG) void k_dot(long n, const double *__restrict__ a,
                                            const double *__restrict__ b, double *out,
                                            double *part, unsigned int *counter)
{
  double v[1] = {0.0}, tot[1];
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) v[0] += a[i] * b[i];
  if (block_reduce_last<1>(v, part, counter, tot) && threadIdx.x == 0) *out = tot[0];
}

__global__ void k_axpby(long n, double a, const double *__restrict__ x, double b,
                        const double *__restrict__ y, double *z)
{
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    z[i] = a * x[i] + b * y[i];
}

// --------------------------------------------------------------------------
// Jacobi PCG (FP64), streamed form
// --------------------------------------------------------------------------
// Partial-sum slots (part = base of the solver's buffer):
//   S  = part                 start kernel, 4 x GS:  (r,z) (r,r) (x,x) (b,b)
//   B  = part + 4 * 65536     update kernel, 4 x GB: (r,z) (r,r) (x,x) 0
//   A  = part + 8 * 65536/2.. SpMV kernel,   1 x GA: (p,q)
#define PART_S(p) (p)                    /* 4 x up to 65536 blocks */
#define PART_B(p) ((p) + 4L * 65536)      /* 4 x up to 16384 blocks */
#define PART_A(p) ((p) + 5L * 65536)      /* 1 x up to 65536 blocks */

// start: r = b - A x; z = dinv r; p = 0; partials (r,z), (r,r), (x,x), (b,b)
template <int BS>
__global__ __launch_bounds__(WG) void k_pcg_start(SellView M, const double *__restrict__ vals,
                                                  const double *__restrict__ b,
                                                  const double *__restrict__ x, double *r,
                                                  double *z, double *p0, double *p1,
                                                  const double *__restrict__ dinv, double *part)
{
  double v[4] = {0, 0, 0, 0};
  double acc[BS];
  long row;
  if (sell_row_split<BS, double, double>(M, vals, x, acc, row)) {
#pragma unroll
    for (int k = 0; k < BS; k++) {
      const long i = row * BS + k;
      const double ri = b[i] - acc[k];
      const double zi = ri * dinv[i];
      r[i] = ri;
      z[i] = zi;
      p0[i] = 0.0;
      p1[i] = 0.0;
      v[0] += ri * zi;
      v[1] += ri * ri;
      v[2] += x[i] * x[i];
      v[3] += b[i] * b[i];
    }
  }
  block_partials<4>(v, PART_S(part));
}

// iteration head shared by the p-update kernels: totals of iteration k's
// residual, convergence test, beta; block 0 records the state
__device__ __forceinline__ bool pcg_head(ugk_state *st, const double *part, unsigned GS, unsigned GB,
                                         long &k, double &beta)
{
  k = st->iter;
  double t[4];
  if (k == 0) reduce_prev<4>(PART_S(part), GS, t);
  else reduce_prev<4>(PART_B(part), GB, t);
  const double bb = (k == 0) ? t[3] : st->bb;
  const int done = pcg_test(st, k, t[1], t[2], bb, t[0]);
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  beta = (k == 0) ? 0.0 : t[0] / st->rz[(k + 1) & 1];
  if (lead) {
    st->rr = t[1];
    st->xx = t[2];
    if (k == 0) st->bb = bb;
    if (done) st->done = done;
    else {
      st->rz[k & 1] = t[0];
      st->rzB = t[0];
      st->beta = beta;
      st->pidx = k & 1;
    }
  }
  return done != 0;
}

// variant 1, kernel 1: p = z + beta p (in place)
__global__ __launch_bounds__(WG) void k_pcg_p(long n, const double *__restrict__ z, double *p,
                                              ugk_state *st, const double *part, unsigned GS,
                                              unsigned GB)
{
  if (st->done) return;
  long k;
  double beta;
  if (pcg_head(st, part, GS, GB, k, beta)) return;
  long i1_, i2_;
  xcd_chunk(n, i1_, i2_);
  for (long i = i1_ + threadIdx.x; i < i2_; i += WG)
    p[i] = z[i] + beta * p[i];
}

// variant 1, kernel 2: q = A p, partial (p, q); A's values in FP64 or, for
// the two-level mixed form, from the FP32 copy (vectors and sums in FP64)
template <int BS, typename T = double>
__global__ __launch_bounds__(WG) void k_pcg_q(SellView M, const T *__restrict__ vals,
                                              const double *__restrict__ p, double *__restrict__ q,
                                              ugk_state *st, double *part)
{
  if (st->done) return;
  double v[1] = {0.0};
  double acc[BS];
  long row;
  if (sell_row_split<BS, T, double>(M, vals, p, acc, row)) {
#pragma unroll
    for (int r = 0; r < BS; r++) {
      const long i = row * BS + r;
      q[i] = acc[r];
      v[0] += p[i] * acc[r];
    }
  }
  block_partials<1>(v, PART_A(part));
}

// variant 0, the single kernel before the update: head + p' = z + beta p
// (own rows written to the other ping-pong buffer) + q = A p' with the
// neighbours' p' formed on the fly from gathered z and p
template <int BS>
__global__ __launch_bounds__(WG) void k_pcg_a(SellView M, const double *__restrict__ vals,
                                              const double *__restrict__ z, double *p0, double *p1,
                                              double *__restrict__ q, ugk_state *st, double *part,
                                              unsigned GS, unsigned GB)
{
  if (st->done) return;
  long k;
  double beta;
  if (pcg_head(st, part, GS, GB, k, beta)) return;
  const double *pold = (k & 1) ? p0 : p1;
  double *pnew = (k & 1) ? p1 : p0;
  const long nb = gridDim.x;
  const long lb = remap_block(blockIdx.x, nb);
  const long sl = lb * NW + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  double v[1] = {0.0};
  if (sl < M.n_slices) {
    const int L = M.slen[sl];
    const long so = M.soff[sl];
    const int *cp = M.cols + so * 64 + lane;
    double acc[BS];
#pragma unroll
    for (int r = 0; r < BS; r++) acc[r] = 0.0;
#pragma unroll 3
    for (int s = 0; s < L; s++) {
      const long J = cp[(long)s * 64];
      double xv[BS];
#pragma unroll
      for (int c = 0; c < BS; c++) xv[c] = z[J * BS + c] + beta * pold[J * BS + c];
      double vv[BS * BS];
      ld_block<BS, false>(vals, so + s, lane, vv);
#pragma unroll
      for (int r = 0; r < BS; r++)
#pragma unroll
        for (int c = 0; c < BS; c++) acc[r] += vv[r * BS + c] * xv[c];
    }
    const long row = sl * 64 + lane;
    if (row < M.n_brows) {
#pragma unroll
      for (int r = 0; r < BS; r++) {
        const long i = row * BS + r;
        const double pn = z[i] + beta * pold[i];
        pnew[i] = pn;
        q[i] = acc[r];
        v[0] += pn * acc[r];
      }
    }
  }
  block_partials<1>(v, PART_A(part));
}

// ---- single-reduction Jacobi PCG (Chronopoulos-Gear) ----------------------
// u = dinv r, w = A u, gamma = (r,u), delta = (w,u);
// beta_i = gamma_i / gamma_{i-1}, alpha_i = gamma_i / (delta_i - beta_i gamma_i / alpha_{i-1});
// p = u + beta p, s = w + beta s (s = A p), x += alpha p, r -= alpha s.
// PART_B holds two parity slots of (gamma, rr, xx, bb) partials.
#define PART_B2(p, par) ((par) ? (p) + 7L * 65536 : PART_B(p))   /* 4 x up to 16384 blocks each */

template <int BS>
__global__ __launch_bounds__(WG) void k_cg1_start(SellView M, const double *__restrict__ vals,
                                                  const double *__restrict__ b,
                                                  const double *__restrict__ x, double *r, double *u,
                                                  double *p, double *s,
                                                  const double *__restrict__ dinv, double *part)
{
  double v[4] = {0, 0, 0, 0};
  double acc[BS];
  long row;
  if (sell_row_split<BS, double, double>(M, vals, x, acc, row)) {
#pragma unroll
    for (int k = 0; k < BS; k++) {
      const long i = row * BS + k;
      const double ri = b[i] - acc[k];
      const double ui = ri * dinv[i];
      r[i] = ri;
      u[i] = ui;
      p[i] = 0.0;
      s[i] = 0.0;
      v[0] += ri * ui;
      v[1] += ri * ri;
      v[2] += x[i] * x[i];
      v[3] += b[i] * b[i];
    }
  }
  // grid of this kernel may exceed the parity slot's 16384 blocks: the slot
  // is sized for vec grids; the start is launched on a vec-sized grid below
  block_partials<4>(v, PART_B2(part, 0));
}

// w = A u, partial (w, u); latches the iteration for the vector kernel
template <int BS>
__global__ __launch_bounds__(WG) void k_cg1_w(SellView M, const double *__restrict__ vals,
                                              const double *__restrict__ u, double *__restrict__ w,
                                              ugk_state *st, double *part)
{
  if (st->done) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) st->cur = st->iter;
  double v[1] = {0.0};
  double acc[BS];
  long row;
  if (sell_row_split<BS, double, double>(M, vals, u, acc, row)) {
#pragma unroll
    for (int r = 0; r < BS; r++) {
      const long i = row * BS + r;
      w[i] = acc[r];
      v[0] += u[i] * acc[r];
    }
  }
  block_partials<1>(v, PART_A(part));
}

__global__ __launch_bounds__(WG) void k_cg1_upd(long n, double *__restrict__ x, double *__restrict__ r,
                                                double *__restrict__ u, double *__restrict__ p,
                                                double *__restrict__ s, const double *__restrict__ w,
                                                const double *__restrict__ dinv, ugk_state *st,
                                                double *part, unsigned G0, unsigned GB, unsigned GA)
{
  if (st->done) return;
  const long i0 = st->cur;
  double t[4], d1[1];
  reduce_prev<4>(PART_B2(part, i0 & 1), i0 == 0 ? G0 : GB, t);
  reduce_prev<1>(PART_A(part), GA, d1);
  const double gam = t[0], dlt = d1[0];
  const double bb = (i0 == 0) ? t[3] : st->bb;
  int done = pcg_test(st, i0, t[1], t[2], bb, gam);
  const double beta = (i0 == 0) ? 0.0 : gam / st->gam[(i0 + 1) & 1];
  const double den = (i0 == 0) ? dlt : dlt - beta * gam / st->alf[(i0 + 1) & 1];
  const double alpha = gam / den;
  if (!done && (!(den > 0.0) || !(alpha == alpha))) done = 2;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (lead) {
    st->rr = t[1];
    st->xx = t[2];
    if (i0 == 0) st->bb = bb;
    if (done) st->done = done;
    else {
      st->gam[i0 & 1] = gam;
      st->alf[i0 & 1] = alpha;
      st->alpha = alpha;
      st->beta = beta;
    }
  }
  if (done) return;
  double v[4] = {0, 0, 0, 0};
  long i1, i2;
  xcd_chunk_even(n, i1, i2);
  // two entries per lane: 16-byte loads and stores
  for (long i = i1 + 2 * threadIdx.x; i < i2; i += 2 * WG) {
    if (i + 1 < i2) {
      const dbl2 u2 = *(const dbl2 *)(u + i), p2 = *(const dbl2 *)(p + i), w2 = *(const dbl2 *)(w + i),
                 s2 = *(const dbl2 *)(s + i), x2 = *(const dbl2 *)(x + i), r2 = *(const dbl2 *)(r + i),
                 d2 = *(const dbl2 *)(dinv + i);
      const dbl2 pi = u2 + beta * p2, si = w2 + beta * s2;
      const dbl2 xi = x2 + alpha * pi, ri = r2 - alpha * si;
      const dbl2 ui = ri * d2;
      *(dbl2 *)(p + i) = pi;
      *(dbl2 *)(s + i) = si;
      *(dbl2 *)(x + i) = xi;
      *(dbl2 *)(r + i) = ri;
      *(dbl2 *)(u + i) = ui;
      v[0] += ri.x * ui.x + ri.y * ui.y;
      v[1] += ri.x * ri.x + ri.y * ri.y;
      v[2] += xi.x * xi.x + xi.y * xi.y;
    } else {
      const double pi = u[i] + beta * p[i];
      const double si = w[i] + beta * s[i];
      const double xi = x[i] + alpha * pi;
      const double ri = r[i] - alpha * si;
      const double ui = ri * dinv[i];
      p[i] = pi;
      s[i] = si;
      x[i] = xi;
      r[i] = ri;
      u[i] = ui;
      v[0] += ri * ui;
      v[1] += ri * ri;
      v[2] += xi * xi;
    }
  }
  block_partials<4>(v, PART_B2(part, (i0 + 1) & 1));
  if (lead) st->iter = i0 + 1;
}

// update: alpha = rz / (p, q); x += alpha p; r -= alpha q; z = dinv r;
// partials (r,z), (r,r), (x,x); block 0 advances the iteration count
__global__ __launch_bounds__(WG) void k_pcg_b(long n, double *__restrict__ x, double *__restrict__ r,
                                              double *__restrict__ z, const double *p0,
                                              const double *p1, const double *__restrict__ q,
                                              const double *__restrict__ dinv, ugk_state *st,
                                              double *part, unsigned GA, int pingpong)
{
  if (st->done) return;
  double t[1];
  reduce_prev<1>(PART_A(part), GA, t);
  const double pq = t[0];
  const double alpha = st->rzB / pq;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (!(pq > 0.0) || !(alpha == alpha)) {
    if (lead) st->done = 2;
    return;
  }
  const double *p = (pingpong && st->pidx) ? p1 : p0;
  double v[4] = {0, 0, 0, 0};
  long i1_, i2_;
  xcd_chunk(n, i1_, i2_);
  for (long i = i1_ + threadIdx.x; i < i2_; i += WG) {
    const double xi = x[i] + alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    const double zi = ri * dinv[i];
    x[i] = xi;
    r[i] = ri;
    z[i] = zi;
    v[0] += ri * zi;
    v[1] += ri * ri;
    v[2] += xi * xi;
  }
  block_partials<4>(v, PART_B(part));
  if (lead) {
    st->alpha = alpha;
    st->iter = st->iter + 1;
  }
}

// Sense-counting grid barrier for a co-resident (cooperative) grid: one
// ticket per block on `bar[0]`; the last arrival resets it and bumps the
// generation `bar[1]` that the others wait on.  Every wait ends because the
// launch is cooperative (all blocks resident) and every block arrives.
// The wait is bounded by `limit` device wall-clock ticks: a block that gives
// up raises bar[2], which every later wait sees at once, and returns false.
__device__ __forceinline__ bool grid_barrier(unsigned int *bar, unsigned G, long limit)
{
  __shared__ int s_ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    const unsigned gen = __hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned t = __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (t == G - 1) {
      __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&bar[1], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      const long t0 = (long)wall_clock64();
      unsigned spins = 0;
      while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
        __builtin_amdgcn_s_sleep(1);
        if ((++spins & 31u) == 0u &&
            (__hip_atomic_load(&bar[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u ||
             (long)wall_clock64() - t0 > limit)) {
          __hip_atomic_store(&bar[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ok = 0;
          break;
        }
      }
    }
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

// whole PCG in one cooperative launch; same arithmetic as k_pcg_a/k_pcg_b
// but with the two reductions finished redundantly in every block after a
// grid barrier.  Every block takes the same decisions, so all leave together.
template <int BS, bool CGSYNC>
__global__ __launch_bounds__(WG) void k_pcg_persistent(SellView M, const double *__restrict__ vals,
                                                       long n, double *x, double *r, double *z,
                                                       double *p0, double *p1, double *q,
                                                       const double *__restrict__ dinv,
                                                       ugk_state *st, double *part, unsigned GS, long limit)
{
  cg::grid_group grid = cg::this_grid();
  const unsigned G = gridDim.x;
  double *partA = PART_B(part), *partB = PART_B(part) + G;   // after the start partials
  double t0[4];
  reduce_prev<4>(PART_S(part), GS, t0);      // written by the start kernel
  long it = 0;
  double rz = t0[0];
  double beta = 0.0;
  const double nbb = sqrt(t0[3]), normA = st->normA, tol = st->tol;
  const int rule = st->rule;
  const long maxit = st->maxit;
  int done = pcg_test(st, 0, t0[1], t0[2], t0[3], t0[0]);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double rr = t0[1], xx = t0[2], alpha = 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) st->bb = t0[3];
  while (!done) {
    const double *pold = (it & 1) ? p1 : p0;
    double *pnew = (it & 1) ? p0 : p1;
    double v[1] = {0.0};
    for (long sl = (long)blockIdx.x * NW + wave; sl < M.n_slices; sl += (long)G * NW) {
      const int L = M.slen[sl];
      const long so = M.soff[sl];
      const int *cp = M.cols + so * 64 + lane;
      double acc[BS];
#pragma unroll
      for (int rr2 = 0; rr2 < BS; rr2++) acc[rr2] = 0.0;
#pragma unroll 3
      for (int s = 0; s < L; s++) {
        const long J = cp[(long)s * 64];
        double xv[BS];
#pragma unroll
        for (int c = 0; c < BS; c++) xv[c] = z[J * BS + c] + beta * pold[J * BS + c];
        double vv[BS * BS];
        ld_block<BS, false>(vals, so + s, lane, vv);
#pragma unroll
        for (int r2 = 0; r2 < BS; r2++)
#pragma unroll
          for (int c = 0; c < BS; c++) acc[r2] += vv[r2 * BS + c] * xv[c];
      }
      const long row = sl * 64 + lane;
      if (row < M.n_brows) {
#pragma unroll
        for (int r2 = 0; r2 < BS; r2++) {
          const long i = row * BS + r2;
          const double pn = z[i] + beta * pold[i];
          pnew[i] = pn;
          q[i] = acc[r2];
          v[0] += pn * acc[r2];
        }
      }
    }
    block_publish<1>(v, partA);
    if (CGSYNC) grid.sync();
    else if (!grid_barrier(&st->counter[6], G, limit)) {
      if (threadIdx.x == 0) {          // a block missing from the barrier: end with an error
        st->err = 5;
        st->done = 2;
      }
      return;
    }
    double tA[1];
    grid_partials_sum<1>(partA, G, tA);
    const double pq = tA[0];
    alpha = rz / pq;
    double w[3] = {0, 0, 0};
    for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)G * WG) {
      const double xi = x[i] + alpha * pnew[i];
      const double ri = r[i] - alpha * q[i];
      const double zi = ri * dinv[i];
      x[i] = xi;
      r[i] = ri;
      z[i] = zi;
      w[0] += ri * zi;
      w[1] += ri * ri;
      w[2] += xi * xi;
    }
    block_publish<3>(w, partB);
    if (CGSYNC) grid.sync();
    else if (!grid_barrier(&st->counter[6], G, limit)) {
      if (threadIdx.x == 0) {          // a block missing from the barrier: end with an error
        st->err = 5;
        st->done = 2;
      }
      return;
    }
    double tB[3];
    grid_partials_sum<3>(partB, G, tB);
    it++;
    rr = tB[1];
    xx = tB[2];
    beta = tB[0] / rz;
    const double nr = sqrt(rr);
    if (!(pq > 0.0) || !(tB[0] == tB[0])) done = 2;
    else if (rule == 1 ? nr <= tol * nbb : nr <= tol * (normA * sqrt(xx) + nbb)) done = 1;
    else if (it >= maxit) done = 3;
    rz = tB[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) st->rz[it & 1] = rz;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->iter = it;
    st->rr = rr;
    st->xx = xx;
    st->beta = beta;
    st->alpha = alpha;
    st->done = done;
  }
  (void)partB;
}

// --------------------------------------------------------------------------
// Mixed precision: FP32 CG (K in FP32, vectors in FP32, dot products summed
// in FP64) that is never restarted; whenever its recursive residual has
// dropped by `delta` since the last true residual, the FP32 solution
// increment is folded into the FP64 solution and r = b - K x is recomputed
// with the FP64 matrix (a "reliable update").  Convergence is only declared
// on such an FP64 residual.
// --------------------------------------------------------------------------
#define PART_R(p) ((p) + 6L * 65536)      /* 2 x up to 65536 blocks */

template <int BS>
__global__ __launch_bounds__(WG) void k_mx_start(SellView M, const double *__restrict__ vals,
                                                 const double *__restrict__ b,
                                                 const double *__restrict__ x64, double *r64,
                                                 float *r32, float *z32, float *p32, float *x32,
                                                 const float *__restrict__ dinv, double *part)
{
  double v[4] = {0, 0, 0, 0};
  double acc[BS];
  long row;
  if (sell_row_split<BS, double, double>(M, vals, x64, acc, row)) {
#pragma unroll
    for (int k = 0; k < BS; k++) {
      const long i = row * BS + k;
      const double ri = b[i] - acc[k];
      const float rf = (float)ri;
      const float zf = rf * dinv[i];
      r64[i] = ri;
      r32[i] = rf;
      z32[i] = zf;
      p32[i] = 0.0f;
      x32[i] = 0.0f;
      v[0] += (double)rf * (double)zf;
      v[1] += ri * ri;
      v[3] += b[i] * b[i];
    }
  }
  block_partials<4>(v, PART_S(part));
}

// head: totals of iteration k's residual -- the start, a refresh (true FP64
// residual) or the FP32 update -- convergence on true residuals only, beta,
// p = z + beta p
__global__ __launch_bounds__(WG) void k_mx_p(long n, const float *__restrict__ z, float *p,
                                             ugk_state *st, const double *part, unsigned GS,
                                             unsigned GB, unsigned GR)
{
  if (st->done) return;
  const long k = st->iter;
  double t[4];
  const bool fresh = (k == 0) || (st->ru_at == k);
  if (k == 0) reduce_prev<4>(PART_S(part), GS, t);
  else if (fresh) { double u[2]; reduce_prev<2>(PART_R(part), GR, u); t[0] = u[0]; t[1] = u[1]; t[2] = t[3] = 0.0; }
  else { double u[2]; reduce_prev<2>(PART_B(part), GB, u); t[0] = u[0]; t[1] = u[1]; t[2] = t[3] = 0.0; }
  const double bb = (k == 0) ? t[3] : st->bb;
  const double nb = sqrt(bb), nr = sqrt(t[1]);
  int done = 0;
  // a refresh that does not improve the best true residual by 10% is a
  // stall; five in a row end the solve (stagnation at the FP32 floor)
  const bool progress = (k == 0) || t[1] < 0.81 * st->best_true;
  const long stall = fresh ? (progress ? 0 : st->n_stall + 1) : st->n_stall;
  if (!(nb > 0.0)) done = 1;
  else if (!(t[0] == t[0]) || !(t[1] == t[1])) done = 2;
  else if (fresh && nr <= st->tol * nb) done = 1;
  else if (stall >= 5) done = 2;
  else if (k >= st->maxit) done = 3;
  // classical refinement (st->restart): each refresh starts a new inner CG
  // on the correction equation, p = z
  const double beta = (k == 0 || (fresh && st->restart)) ? 0.0 : t[0] / st->rz[(k + 1) & 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (k == 0) st->bb = bb;
    if (fresh) {
      st->rr_true = t[1];
      st->maxrr[k & 1] = nr;
      st->n_stall = stall;
      if (progress) st->best_true = t[1];
    }
    st->rr = t[1];
    if (done) st->done = done;
    else { st->rz[k & 1] = t[0]; st->rzB = t[0]; st->beta = beta; }
  }
  if (done) return;
  const float bf = (float)beta;
  long i1_, i2_;
  xcd_chunk(n, i1_, i2_);
  for (long i = i1_ + threadIdx.x; i < i2_; i += WG) p[i] = z[i] + bf * p[i];
}

template <int BS>
__global__ __launch_bounds__(WG) void k_mx_q(SellView M, const float *__restrict__ vals,
                                             const float *__restrict__ p, float *__restrict__ q,
                                             ugk_state *st, double *part)
{
  if (st->done) return;
  double v[1] = {0.0};
  float acc[BS];
  long row;
  if (sell_row_split<BS, float, float>(M, vals, p, acc, row)) {
#pragma unroll
    for (int r = 0; r < BS; r++) {
      const long i = row * BS + r;
      q[i] = acc[r];
      v[0] += (double)p[i] * (double)acc[r];
    }
  }
  block_partials<1>(v, PART_A(part));
}

__global__ __launch_bounds__(WG) void k_mx_b(long n, float *__restrict__ x, float *__restrict__ r,
                                             float *__restrict__ z, const float *__restrict__ p,
                                             const float *__restrict__ q,
                                             const float *__restrict__ dinv, ugk_state *st,
                                             double *part, unsigned GA)
{
  if (st->done) return;
  double t[1];
  reduce_prev<1>(PART_A(part), GA, t);
  const double pq = t[0];
  const double alpha = st->rzB / pq;
  if (!(pq > 0.0) || !(alpha == alpha)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) st->done = 2;
    return;
  }
  const float af = (float)alpha;
  double v[2] = {0, 0};
  long i1_, i2_;
  xcd_chunk(n, i1_, i2_);
  for (long i = i1_ + threadIdx.x; i < i2_; i += WG) {
    const float xi = x[i] + af * p[i];
    const float ri = r[i] - af * q[i];
    const float zi = ri * dinv[i];
    x[i] = xi;
    r[i] = ri;
    z[i] = zi;
    v[0] += (double)ri * (double)zi;
    v[1] += (double)ri * (double)ri;
  }
  block_partials<2>(v, PART_B(part));
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->alpha = alpha;
    st->iter = st->iter + 1;
  }
}

// refresh test on the FP32 residual just produced; on refresh fold x32 into
// x64 (the FP64 residual follows in k_mx_r)
__global__ __launch_bounds__(WG) void k_mx_ru(long n, double *x64, float *x32, ugk_state *st,
                                              const double *part, unsigned GB)
{
  if (st->done) return;
  const long k1 = st->iter;
  double t[2];
  reduce_prev<2>(PART_B(part), GB, t);
  const double rn = sqrt(t[1]);
  const double mx = st->maxrr[(k1 + 1) & 1];
  const bool refresh = !(rn >= st->delta * mx) || rn <= st->tol * sqrt(st->bb) || k1 >= st->maxit ||
                       k1 - st->ru_at >= st->max_gap;
  if (!refresh) {
    if (blockIdx.x == 0 && threadIdx.x == 0) st->maxrr[k1 & 1] = fmax(mx, rn);
    return;
  }
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) {
    x64[i] += (double)x32[i];
    x32[i] = 0.0f;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) st->ru_req = k1;
}

template <int BS>
__global__ __launch_bounds__(WG) void k_mx_r(SellView M, const double *__restrict__ vals,
                                             const double *__restrict__ b,
                                             const double *__restrict__ x64, double *r64,
                                             float *r32, float *z32, const float *__restrict__ dinv,
                                             ugk_state *st, double *part)
{
  if (st->done || st->ru_req != st->iter) return;
  double v[2] = {0, 0};
  double acc[BS];
  long row;
  if (sell_row_split<BS, double, double>(M, vals, x64, acc, row)) {
#pragma unroll
    for (int k = 0; k < BS; k++) {
      const long i = row * BS + k;
      const double ri = b[i] - acc[k];
      const float rf = (float)ri;
      const float zf = rf * dinv[i];
      r64[i] = ri;
      r32[i] = rf;
      z32[i] = zf;
      v[0] += (double)rf * (double)zf;
      v[1] += ri * ri;
    }
  }
  block_partials<2>(v, PART_R(part));
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->ru_at = st->iter;
    st->n_ru = st->n_ru + 1;
  }
}

template <int BS>
__global__ __launch_bounds__(WG) void k_residual(SellView M, const double *__restrict__ vals,
                                                 const double *__restrict__ b,
                                                 const double *__restrict__ x, double *r,
                                                 ugk_state *st, double *part)
{
  double v[1] = {0.0}, tot[1];
  double acc[BS];
  long row;
  if (sell_row_split<BS, double, double>(M, vals, x, acc, row)) {
#pragma unroll
    for (int k = 0; k < BS; k++) {
      const long i = row * BS + k;
      const double ri = b[i] - acc[k];
      r[i] = ri;
      v[0] += ri * ri;
    }
  }
  if (block_reduce_last<1>(v, part, &st->counter[4], tot) && threadIdx.x == 0) st->rr = tot[0];
}

// --------------------------------------------------------------------------
// BiCGStab (Jacobi, right preconditioned), five kernels per iteration
// --------------------------------------------------------------------------
// aux[0] = rho (rhat, r) of the current iteration, aux[1] = (rhat, v),
// aux[2] = ||s||^2, aux[3] = (t, s), aux[4] = (t, t), aux[5] = s-converged
__global__ void k_bicgs_init(long n, const double *__restrict__ b, double *r, double *rhat,
                             double *p, double *v, ugk_state *st, double *part)
{
  double w[2] = {0, 0}, tot[2];
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) {
    const double ri = r[i];   // r holds A x on entry
    const double rv = b[i] - ri;
    r[i] = rv;
    rhat[i] = rv;
    p[i] = 0.0;
    v[i] = 0.0;
    w[0] += rv * rv;
    w[1] += b[i] * b[i];
  }
  if (block_reduce_last<2>(w, part, &st->counter[0], tot) && threadIdx.x == 0) {
    st->rr = tot[0];
    st->bb = tot[1];
    st->aux[0] = tot[0];   // rho = (rhat, r) = ||r||^2
    st->rho = 1.0;
    st->alpha = 1.0;
    st->omega = 1.0;
    st->iter = 0;
    const double nb = sqrt(tot[1]);
    int done = 0;
    if (!(nb > 0.0)) done = 1;
    else if (sqrt(tot[0]) <= st->tol * nb) done = 1;
    else if (st->maxit <= 0) done = 3;
    st->done = done;
  }
}

__global__ void k_bicgs_p(long n, const double *__restrict__ r, double *p, const double *__restrict__ v,
                          double *phat, const double *__restrict__ dinv, ugk_state *st)
{
  if (st->done) return;
  const double rho = st->aux[0];
  const double beta = (rho / st->rho) * (st->alpha / st->omega);
  const double om = st->omega;
  const bool first = st->iter == 0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double pi = first ? r[i] : r[i] + beta * (p[i] - om * v[i]);
    p[i] = pi;
    phat[i] = pi * dinv[i];
  }
}

// y = A x and partial dots (y, w1) [and (y, y)] -> aux[slot], aux[slot+1]
template <int BS, int ND>
__global__ __launch_bounds__(WG) void k_spmv_dots(SellView M, const double *__restrict__ vals,
                                                  const double *__restrict__ x, double *y,
                                                  const double *__restrict__ w1, ugk_state *st,
                                                  double *part, int slot, int ctr)
{
  if (st->done) return;
  if (ND == 2 && st->aux[5] != 0.0) return;   // s already converged
  double v[ND], tot[ND];
#pragma unroll
  for (int k = 0; k < ND; k++) v[k] = 0.0;
  double acc[BS];
  long row;
  if (sell_row_split<BS, double, double>(M, vals, x, acc, row)) {
#pragma unroll
    for (int k = 0; k < BS; k++) {
      const long i = row * BS + k;
      y[i] = acc[k];
      v[0] += acc[k] * w1[i];
      if (ND == 2) v[1] += acc[k] * acc[k];
    }
  }
  if (block_reduce_last<ND>(v, part, &st->counter[ctr], tot) && threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < ND; k++) st->aux[slot + k] = tot[k];
  }
}

__global__ __launch_bounds__(WG) void k_bicgs_s(long n, const double *__restrict__ r,
                                                const double *__restrict__ v, double *s,
                                                double *shat, const double *__restrict__ dinv,
                                                ugk_state *st, double *part)
{
  if (st->done) return;
  const double alpha = st->aux[0] / st->aux[1];
  double w[1] = {0.0}, tot[1];
  long i1_, i2_;
  xcd_chunk(n, i1_, i2_);
  for (long i = i1_ + threadIdx.x; i < i2_; i += WG) {
    const double si = r[i] - alpha * v[i];
    s[i] = si;
    shat[i] = si * dinv[i];
    w[0] += si * si;
  }
  if (block_reduce_last<1>(w, part, &st->counter[2], tot) && threadIdx.x == 0) {
    st->aux[2] = tot[0];
    st->alpha = alpha;
    st->aux[5] = (sqrt(tot[0]) <= st->tol * sqrt(st->bb) || !(st->aux[1] != 0.0)) ? 1.0 : 0.0;
  }
}

__global__ __launch_bounds__(WG) void k_bicgs_x(long n, double *x, double *r,
                                                const double *__restrict__ rhat,
                                                const double *__restrict__ phat,
                                                const double *__restrict__ s,
                                                const double *__restrict__ shat,
                                                const double *__restrict__ t, ugk_state *st,
                                                double *part)
{
  if (st->done) return;
  const bool sconv = st->aux[5] != 0.0;
  const double alpha = st->alpha;
  const double omega = sconv ? 0.0 : st->aux[3] / st->aux[4];
  double w[2] = {0, 0}, tot[2];
  long i1_, i2_;
  xcd_chunk(n, i1_, i2_);
  for (long i = i1_ + threadIdx.x; i < i2_; i += WG) {
    const double xi = x[i] + alpha * phat[i] + omega * shat[i];
    const double ri = sconv ? s[i] : s[i] - omega * t[i];
    x[i] = xi;
    r[i] = ri;
    w[0] += rhat[i] * ri;
    w[1] += ri * ri;
  }
  if (block_reduce_last<2>(w, part, &st->counter[5], tot) && threadIdx.x == 0) {
    const long it1 = st->iter + 1;
    st->iter = it1;
    st->rho = st->aux[0];
    st->aux[0] = tot[0];
    st->omega = omega;
    st->rr = tot[1];
    st->aux[5] = 0.0;
    const double nr = sqrt(tot[1]), nbb = sqrt(st->bb);
    int done = 0;
    if (!(tot[1] == tot[1])) done = 2;
    else if (sconv || nr <= st->tol * nbb) done = 1;
    else if (!(omega != 0.0) || !(tot[0] != 0.0)) done = 2;
    else if (it1 >= st->maxit) done = 3;
    st->done = done;
  }
}

// --------------------------------------------------------------------------
// Batched solve: samples of different sizes in one persistent launch.  A
// workgroup owns a sample from start to finish, so all synchronisation is
// workgroup barriers (no grid-wide barriers, no fences); the ticket counter
// deals samples out largest first.  Per iteration: w = K u with (w,u), then
// the single-reduction PCG update of p, s, x, r, u with (r,u), (r,r).
// --------------------------------------------------------------------------
template <int WGS, int NV>
__device__ __forceinline__ void wg_sum(double (&v)[NV], double (*red)[NV])
{
  constexpr int W = WGS / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = wave_sum(v[k]);
    if (lane == 0) red[wave][k] = s;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double a = red[0][k];
    for (int w = 1; w < W; w++) a += red[w][k];
    v[k] = a;
  }
  __syncthreads();
}

// Stop request of a batched launch: the host's flag (mapped host memory,
// written while the kernel runs) or the device flag a workgroup raised when
// one of its waits gave up.
struct BatchStop {
  const int *host;          // may be NULL
  unsigned int *dev;
  long limit;               // ticks of the device wall clock a wait may take
};
__device__ __forceinline__ bool stop_asked(const BatchStop &b)
{
  return (b.host && __hip_atomic_load(b.host, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0) ||
         __hip_atomic_load(b.dev, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
}

// Team of T workgroups on one sample: the slices and vector entries are
// dealt round robin over the team; sums go through per-rank partial slots
// (two buffers, by reduction parity) and a monotonic team counter: barrier k
// waits until the counter reaches T k.  A team forms only from resident
// workgroups (the grid never exceeds residency), so every wait ends; the
// wait is still bounded -- by the wall-clock limit and by a stop request --
// so that a miscount ends the launch with an error instead of hanging it.
// Returns false (in every thread) when the wait gave up.
__device__ __forceinline__ bool team_barrier(unsigned int *ctr, unsigned target, const BatchStop &bs)
{
  __shared__ int s_ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    const long t0 = (long)wall_clock64();
    unsigned spins = 0;
    while (__hip_atomic_load(ctr, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(2);
      if ((++spins & 31u) == 0u && (stop_asked(bs) || (long)wall_clock64() - t0 > bs.limit)) {
        ok = 0;
        __hip_atomic_store(bs.dev, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
    }
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

template <int WGS, int NV>
__device__ __forceinline__ bool team_sum(double (&v)[NV], double (*red)[NV], int T, int rank,
                                         double *tpart, unsigned int *ctr, unsigned &nbar, int &par,
                                         const BatchStop &bs)
{
  wg_sum<WGS, NV>(v, red);
  if (T == 1) return true;
  double *slot = tpart + (size_t)par * 8 * 4;   // [2][8 ranks][4]
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) slot[rank * 4 + k] = v[k];
  }
  nbar++;
  if (!team_barrier(ctr, nbar * (unsigned)T, bs)) return false;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double a = 0.0;
    for (int r = 0; r < T; r++) a += __hip_atomic_load(&slot[r * 4 + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    v[k] = a;
  }
  par ^= 1;
  return true;
}

// the three sums of a block row over its slots for vector v (FP64 or FP32
// values; REM: FP32 plus the bf16 remainder at the same position, summed
// exactly in FP64)
template <typename T, bool REM>
__device__ __forceinline__ void batch_row(const T *__restrict__ vals, const unsigned short *__restrict__ rem,
                                          const double *__restrict__ v, long so, int L,
                                          const int *__restrict__ cp, int lane, double &a0, double &a1, double &a2)
{
  for (int s = 0; s < L; s++) {
    const long J = ldnt(cp + (long)s * 64);
    const double x0 = v[3 * J], x1 = v[3 * J + 1], x2 = v[3 * J + 2];
    T vv[9];
    ld_block<3, true>(vals, so + s, lane, vv);
    double d[9];
#pragma unroll
    for (int q = 0; q < 9; q++) d[q] = (double)vv[q];
    if constexpr (REM) {
      unsigned short u[9];
      ld_block<3, true>(rem, so + s, lane, u);
#pragma unroll
      for (int q = 0; q < 9; q++) d[q] += (double)__uint_as_float((unsigned)u[q] << 16);
    }
    a0 += d[0] * x0 + d[1] * x1 + d[2] * x2;
    a1 += d[3] * x0 + d[4] * x1 + d[5] * x2;
    a2 += d[6] * x0 + d[7] * x1 + d[8] * x2;
  }
}

template <int WGS>
__global__ __launch_bounds__(WGS) void k_batch_pcg(const ugk_bsample *__restrict__ samples,
                                                   const int2 *__restrict__ slots, int nslots,
                                                   unsigned *ticket, unsigned int *tctr,
                                                   double *tpart, ugk_bresult *res, double tol,
                                                   long maxit, BatchStop bs)
{
  constexpr int W = WGS / 64;
  __shared__ int s_k;
  __shared__ double red3[W][3];
  __shared__ double red1[W][1];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (;;) {
    // no new sample once a stop was asked for
    if (threadIdx.x == 0) s_k = stop_asked(bs) ? nslots : (int)atomicAdd(ticket, 1u);
    __syncthreads();
    const int k = s_k;
    __syncthreads();
    if (k >= nslots) return;
    const int id = slots[k].x;            // sample
    const int T = slots[k].y >> 8;        // team size
    const int rank = slots[k].y & 255;    // rank in the team
    const ugk_bsample S = samples[id];
    const long mxit = S.maxit > 0 ? S.maxit : maxit;
    unsigned int *ctr = tctr + id;
    double *tp = tpart + (size_t)id * 2 * 8 * 4;
    unsigned nbar = 0;
    int par = 0;
    const long n = S.n_brows * 3;
    const long e0 = (long)rank * WGS + threadIdx.x, estep = (long)T * WGS;
    const long w0 = (long)rank * W + wave, wstep = (long)T * W;
    long t0 = 0;
    if (threadIdx.x == 0 && rank == 0) t0 = wall_clock64();
    double v3[3] = {0, 0, 0};
    for (long i = e0; i < n; i += estep) {
      const double ri = S.f[i];
      const double ui = ri * S.dinv[i];
      S.x[i] = 0.0;
      S.r[i] = ri;
      S.u[i] = ui;
      S.p[i] = 0.0;
      S.s[i] = 0.0;
      v3[0] += ri * ui;
      v3[1] += ri * ri;
    }
    int status = 0;
    const bool formed = team_sum<WGS, 3>(v3, red3, T, rank, tp, ctr, nbar, par, bs);   // also publishes u
    double gam = v3[0], rr = v3[1], gam_old = 1.0, alf_old = 1.0;
    const double nb = sqrt(rr);
    const bool mixed = S.vals32 != nullptr;
    double rr_ref = rr;        // mixed: ||r||^2 of the last FP64 residual
    double den_old = 0.0;      // (p, s) of the last direction
    long it = 0, it_ref = 0;
    bool exact = true;         // r is an FP64 residual f - K x (the start's is)
    bool refreshed = false;    // r was replaced since the last update
    bool restart = false;      // mixed: the next direction starts afresh (beta = 0)
    if (!formed) status = 4;
    while (status == 0) {
      if ((it & 15) == 15) {           // a stop asked for ends a sample too
        __shared__ int s_stop;
        if (threadIdx.x == 0) s_stop = stop_asked(bs) ? 1 : 0;
        __syncthreads();
        const int stop = s_stop;
        __syncthreads();
        if (stop) { status = 4; break; }
      }
      const bool conv = !(nb > 0.0) || sqrt(rr) <= tol * nb;
      if (mixed && !exact && (conv || restart || rr < 0.01 * rr_ref || it - it_ref >= 500)) {
        // FP64 refresh: r = f - K x, u = dinv r, (r, u), (r, r); the
        // direction is kept (reliable update)
        double v3r[3] = {0, 0, 0};
        for (long sl = w0; sl < S.n_slices; sl += wstep) {
          const long so = S.soff[sl];
          double a[3] = {0.0, 0.0, 0.0};
          batch_row<double, false>(S.vals, nullptr, S.x, so, S.slen[sl], S.cols + so * 64 + lane, lane, a[0], a[1],
                                   a[2]);
          const long row = sl * 64 + lane;
          if (row < S.n_brows) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
              const long i = 3 * row + c;
              const double ri = S.f[i] - a[c], ui = ri * S.dinv[i];
              S.r[i] = ri;
              S.u[i] = ui;
              v3r[0] += ri * ui;
              v3r[1] += ri * ri;
            }
          }
        }
        if (!team_sum<WGS, 3>(v3r, red3, T, rank, tp, ctr, nbar, par, bs)) { status = 4; break; }   // publishes r, u
        gam = v3r[0];
        rr = v3r[1];
        rr_ref = rr;
        it_ref = it;
        exact = true;
        refreshed = true;
        continue;
      }
      if (conv) { status = 1; break; }
      if (!(gam == gam) || !(rr == rr)) { status = 2; break; }
      if (it >= mxit) { status = 3; break; }
      double v3s[3] = {0.0, 0.0, 0.0};   // (w, u) [, (u, s) after a refresh]
      for (long sl = w0; sl < S.n_slices; sl += wstep) {
        const int L = S.slen[sl];
        const long so = S.soff[sl];
        const int *cp = S.cols + so * 64 + lane;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        if (!mixed) batch_row<double, false>(S.vals, nullptr, S.u, so, L, cp, lane, a0, a1, a2);
        else if (S.rem16) batch_row<float, true>(S.vals32, S.rem16, S.u, so, L, cp, lane, a0, a1, a2);
        else batch_row<float, false>(S.vals32, nullptr, S.u, so, L, cp, lane, a0, a1, a2);
        const long row = sl * 64 + lane;
        if (row < S.n_brows) {
          const double u0 = S.u[3 * row], u1 = S.u[3 * row + 1], u2 = S.u[3 * row + 2];
          if (mixed) {   // the FP32 copy's row-sum correction: rigid translations stay exact
            const float *C = S.corr32;
            const long nbr = S.n_brows;
            a0 += (double)C[row] * u0 + (double)C[nbr + row] * u1 + (double)C[2 * nbr + row] * u2;
            a1 += (double)C[3 * nbr + row] * u0 + (double)C[4 * nbr + row] * u1 + (double)C[5 * nbr + row] * u2;
            a2 += (double)C[6 * nbr + row] * u0 + (double)C[7 * nbr + row] * u1 + (double)C[8 * nbr + row] * u2;
          }
          S.w[3 * row] = a0;
          S.w[3 * row + 1] = a1;
          S.w[3 * row + 2] = a2;
          v3s[0] += a0 * u0 + a1 * u1 + a2 * u2;
          if (refreshed) v3s[1] += u0 * S.s[3 * row] + u1 * S.s[3 * row + 1] + u2 * S.s[3 * row + 2];
        }
      }
      if (refreshed) {
        if (!team_sum<WGS, 3>(v3s, red3, T, rank, tp, ctr, nbar, par, bs)) { status = 4; break; }   // publishes w
      } else {
        double v1[1] = {v3s[0]};
        if (!team_sum<WGS, 1>(v1, red1, T, rank, tp, ctr, nbar, par, bs)) { status = 4; break; }
        v3s[0] = v1[0];
      }
      const double dlt = v3s[0];
      const bool fresh = it == 0 || restart;
      const double beta = fresh ? 0.0 : gam / gam_old;
      // (p, K p) of p = u + beta p_old: after a refresh from its three terms
      // ((u, s_old) is no longer -gam / alpha_old), else the single-reduction form
      const double den = fresh ? dlt
                         : (refreshed ? dlt + 2.0 * beta * v3s[1] + beta * beta * den_old : dlt - beta * gam / alf_old);
      const double alpha = gam / den;
      if (!(den > 0.0) || !(alpha == alpha)) {
        // mixed: FP32 rounding can break the short recurrence near
        // stagnation; restart along u from an FP64 residual, once
        if (mixed && !fresh) {
          restart = true;
          exact = false;
          continue;
        }
        status = 2;
        break;
      }
      double w3[3] = {0, 0, 0};
      for (long i = e0; i < n; i += estep) {
        const double pi = S.u[i] + beta * S.p[i];
        const double si = S.w[i] + beta * S.s[i];
        const double ri = S.r[i] - alpha * si;
        const double ui = ri * S.dinv[i];
        S.p[i] = pi;
        S.s[i] = si;
        S.x[i] += alpha * pi;
        S.r[i] = ri;
        S.u[i] = ui;
        w3[0] += ri * ui;
        w3[1] += ri * ri;
      }
      if (!team_sum<WGS, 3>(w3, red3, T, rank, tp, ctr, nbar, par, bs)) { status = 4; break; }   // publishes u
      gam_old = gam;
      alf_old = alpha;
      den_old = den;
      gam = w3[0];
      rr = w3[1];
      it++;
      exact = false;
      refreshed = false;
      restart = false;
    }
    if (threadIdx.x == 0 && rank == 0) {
      ugk_bresult R;
      R.iters = it;
      R.relres = nb > 0.0 ? sqrt(rr) / nb : 0.0;
      R.status = status;
      R.team = T;
      R.t_start = t0;
      R.t_end = wall_clock64();
      res[id] = R;
    }
    __syncthreads();
    if (status == 4) return;           // this workgroup takes no further sample
  }
}

// ==========================================================================
// C-ABI
// ==========================================================================
// --------------------------------------------------------------------------
// flexible CG with the learned preconditioner z = M(r), truncated to the
// last m directions (Notay's FCG(m)):
//   z    = dinv r + s mask sqrt(dinv) N(sqrt(dinv) r / s),  s = rms(sqrt(dinv) r)
//   p_k  = z_k - sum_j (z_k, q_j) / (p_j, q_j) p_j    over the m previous
//          directions p_j, q_j = A p_j (for m = 1 the Polak-Ribiere update
//          (z_k, r_k - r_{k-1}) / (z_{k-1}, r_{k-1}), since r_k - r_{k-1} =
//          -alpha_{k-1} q_{k-1})
//   alpha = (p_k, r_k) / (p_k, q_k)
// Direction k lives in ring slot k mod m of P and Q (m vectors each).  When
// (p_k, r_k) <= 0 (not a descent direction) or is not finite, the iteration
// restarts along dinv r and stays Jacobi PCG from there on (st->fb).
// Partials: B2(parity) update -> head: (r, dinv r), (r,r), (x,x), (b,b);
// S combine -> p: (z,r), (dinv r, r), (z, q_j) x 8, (p_j, r) x 8;
// A SpMV -> update: (p,q).
// --------------------------------------------------------------------------
#define FCG_MAXM 8
#define FCG_NV (2 + 2 * FCG_MAXM)

template <int BS>
__global__ __launch_bounds__(WG) void k_fcg_start(SellView M, const double *__restrict__ vals,
                                                  const double *__restrict__ b,
                                                  const double *__restrict__ x, double *r,
                                                  const double *__restrict__ dinv, double *part)
{
  double v[4] = {0, 0, 0, 0};
  double acc[BS];
  long row;
  if (sell_row_split<BS, double, double>(M, vals, x, acc, row)) {
#pragma unroll
    for (int k = 0; k < BS; k++) {
      const long i = row * BS + k;
      const double ri = b[i] - acc[k];
      r[i] = ri;
      v[0] += dinv[i] * ri * ri;
      v[1] += ri * ri;
      v[2] += x[i] * x[i];
      v[3] += b[i] * b[i];
    }
  }
  block_partials<4>(v, PART_B2(part, 0));
}

__global__ __launch_bounds__(WG) void k_fcg_head(long n, const double *__restrict__ r,
                                                 const double *__restrict__ dinv, unsigned short *x0,
                                                 ugk_state *st, const double *part, unsigned G0, unsigned GB)
{
  if (st->done) return;
  const long k = st->iter;
  // mixed form (max_gap > 0): the residual of this iteration is the FP64
  // b - K x when the refresh kernel ran for it (its partials, G0 blocks)
  const bool mixed = st->max_gap > 0;
  const bool fresh = k == 0 || (mixed && st->ru_at == k);
  double t[4];
  reduce_prev<4>(PART_B2(part, k & 1), fresh ? G0 : GB, t);
  const double bb = (k == 0) ? t[3] : st->bb;
  int done = pcg_test(st, k, t[1], t[2], bb, t[0]);
  long req = -1;
  if (mixed && !fresh) {
    // convergence is only accepted on a recomputed residual
    if (done == 1) {
      done = 0;
      req = k + 1;
    } else if (done == 0 && (t[1] < st->delta * st->delta * st->rr_true || k - st->ru_at >= st->max_gap))
      req = k + 1;
  }
  const int fb = st->fb;
  const double sc = sqrt(t[0] / (double)n);
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (lead) {
    st->cur = k;
    st->rr = t[1];
    st->xx = t[2];
    if (k == 0) st->bb = bb;
    if (mixed && fresh) st->rr_true = t[1];
    st->aux[6] = t[0];   // (r, dinv r) of this iteration, for the Jacobi direction update
    if (req >= 0) st->ru_req = req;
    if (done) st->done = done;
    st->halt = (done || fb) ? 1 : 0;
    st->net_s = sc;
  }
  if (done || fb || !x0) return;
  const double inv = sc > 0.0 ? 1.0 / sc : 0.0;
  // split format, one voxel (3 DOFs) per thread: hi[4], mid[4], lo[4] of its
  // exact truncation split, channel 3 zero
  typedef unsigned short us4 __attribute__((ext_vector_type(4)));
  for (long v = (long)blockIdx.x * WG + threadIdx.x; v < n / 3; v += (long)gridDim.x * WG) {
    us4 part[3] = {(us4){0, 0, 0, 0}, (us4){0, 0, 0, 0}, (us4){0, 0, 0, 0}};
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const long i = 3 * v + c;
      const float f = (float)(sqrt(dinv[i]) * r[i] * inv);
      const unsigned u = __float_as_uint(f), h = u & 0xffff0000u;
      const float r1 = f - __uint_as_float(h);
      const unsigned m = __float_as_uint(r1) & 0xffff0000u;
      const float r2 = r1 - __uint_as_float(m);
      part[0][c] = (unsigned short)(h >> 16);
      part[1][c] = (unsigned short)(m >> 16);
      part[2][c] = (unsigned short)(__float_as_uint(r2) >> 16);
    }
    us4 *o = (us4 *)(x0 + v * 12);
    o[0] = part[0];
    o[1] = part[1];
    o[2] = part[2];
  }
}

__global__ __launch_bounds__(WG) void k_fcg_combine(long n, const double *__restrict__ r,
                                                    const double *__restrict__ dinv,
                                                    const float *__restrict__ mask,
                                                    const float *__restrict__ out,
                                                    const double *__restrict__ P,
                                                    const double *__restrict__ Q, int m,
                                                    const double *__restrict__ zc,
                                                    double *__restrict__ z, const ugk_state *st,
                                                    double *part)
{
  if (st->done) return;
  const long k = st->cur;
  const int nv = k < m ? (int)k : m;          // slots 0 .. nv-1 hold directions
  const double sc = st->net_s;
  const int fb = st->fb;
  double v[FCG_NV];
#pragma unroll
  for (int j = 0; j < FCG_NV; j++) v[j] = 0.0;
  long i1_, i2_;
  xcd_chunk(n, i1_, i2_);
  for (long i = i1_ + threadIdx.x; i < i2_; i += WG) {
    const double ri = r[i];
    const double a = dinv[i] * ri;
    double zi = a;
    if (!fb && out) {
      const long vx = i / 3;
      zi += sc * (double)mask[i] * sqrt(dinv[i]) * (double)out[vx * 4 + (i - 3 * vx)];
    }
    if (zc) zi += zc[i];
    z[i] = zi;
    v[0] += zi * ri;
    v[1] += a * ri;
#pragma unroll
    for (int j = 0; j < FCG_MAXM; j++) {
      if (j < nv) {
        v[2 + j] += zi * Q[(long)j * n + i];
        v[2 + FCG_MAXM + j] += P[(long)j * n + i] * ri;
      }
    }
  }
  block_partials<FCG_NV>(v, PART_S(part));
}

__global__ __launch_bounds__(WG) void k_fcg_p(long n, const double *__restrict__ r,
                                              const double *__restrict__ dinv,
                                              const double *__restrict__ z, double *P, int m,
                                              ugk_state *st, const double *part, unsigned GV)
{
  if (st->done) return;
  const long k = st->cur;
  const int nv = k < m ? (int)k : m;
  double t[FCG_NV];
  reduce_prev<FCG_NV>(PART_S(part), GV, t);
  double beta[FCG_MAXM];
  double pr = t[0];
  bool ok = true;
#pragma unroll
  for (int j = 0; j < FCG_MAXM; j++) {
    beta[j] = (j < nv) ? -t[2 + j] / st->pqs[j] : 0.0;
    ok = ok && isfinite(beta[j]);
    pr += beta[j] * t[2 + FCG_MAXM + j];
  }
  // st->fb is set by block 0 below, so every block reads the copy the head
  // kernel latched (halt == fb here, since done returned above)
  const bool restart = !st->halt && !(ok && pr > 0.0 && isfinite(pr));
  if (restart) {
#pragma unroll
    for (int j = 0; j < FCG_MAXM; j++) beta[j] = 0.0;
    pr = t[1];
  }
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (!(pr > 0.0) || !isfinite(pr)) {
    if (lead) {
      st->done = 2;
      st->halt = 1;
    }
    return;
  }
  if (lead) {
    if (restart) {
      st->fb = 1;
      st->fb_iter = k;
    }
    st->rzB = pr;
    st->beta = beta[0];
  }
  const long slot = k % m;
  long i1_, i2_;
  xcd_chunk(n, i1_, i2_);
  for (long i = i1_ + threadIdx.x; i < i2_; i += WG) {
    double acc = restart ? dinv[i] * r[i] : z[i];
#pragma unroll
    for (int j = 0; j < FCG_MAXM; j++)
      if (j < nv) acc += beta[j] * P[(long)j * n + i];   // slot k mod m is read before it is overwritten
    P[slot * n + i] = acc;
  }
}

// q = A p for the direction of this iteration (ring slot), partial (p, q);
// T = float: the FP32 copy of A (mixed form), vectors and sums stay FP64
template <int BS, typename T>
__global__ __launch_bounds__(WG) void k_fcg_q(SellView M, const T *__restrict__ vals,
                                              const double *__restrict__ P, double *__restrict__ Q, long n,
                                              int m, const ugk_state *st, double *part)
{
  if (st->done) return;
  const long slot = st->cur % m;
  const double *p = P + slot * n;
  double *q = Q + slot * n;
  double v[1] = {0.0};
  double acc[BS];
  long row;
  if (sell_row_split<BS, T, double>(M, vals, p, acc, row)) {
#pragma unroll
    for (int r = 0; r < BS; r++) {
      const long i = row * BS + r;
      q[i] = acc[r];
      v[0] += p[i] * acc[r];
    }
  }
  block_partials<1>(v, PART_A(part));
}

// mixed form: when the head kernel of the previous iteration asked for it,
// r = b - A x in FP64 and the head's partials of this iteration taken from it
template <int BS, bool LAZY = false>
__global__ __launch_bounds__(WG) void k_fcg_refresh(SellView M, const double *__restrict__ vals,
                                                    const double *__restrict__ b,
                                                    const double *__restrict__ x, double *r,
                                                    const double *__restrict__ dinv, ugk_state *st,
                                                    double *part)
{
  // the same for every block; LAZY: launched once per chunk, it serves any
  // request made since the last refresh (the head clears it once served)
  if (st->done || (LAZY ? (st->ru_req < 0 || st->ru_req > st->iter) : st->ru_req != st->iter)) return;
  const long k = st->iter;
  double v[4] = {0, 0, 0, 0};
  double acc[BS];
  long row;
  if (sell_row_split<BS, double, double>(M, vals, x, acc, row)) {
#pragma unroll
    for (int c = 0; c < BS; c++) {
      const long i = row * BS + c;
      const double ri = b[i] - acc[c];
      r[i] = ri;
      v[0] += dinv[i] * ri * ri;
      v[1] += ri * ri;
      v[2] += x[i] * x[i];
      v[3] += b[i] * b[i];
    }
  }
  block_partials<4>(v, PART_B2(part, k & 1));
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->ru_at = k;
    st->n_ru += 1;
  }
}

__global__ __launch_bounds__(WG) void k_fcg_update(long n, double *__restrict__ x, double *__restrict__ r,
                                                   const double *__restrict__ P,
                                                   const double *__restrict__ Q, int m,
                                                   const double *__restrict__ dinv, ugk_state *st,
                                                   double *part, unsigned GA)
{
  if (st->done) return;
  const long k = st->cur;
  const long slot = k % m;
  const double *p = P + slot * n, *q = Q + slot * n;
  double t[1];
  reduce_prev<1>(PART_A(part), GA, t);
  const double pq = t[0];
  const double alpha = st->rzB / pq;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (!(pq > 0.0) || !isfinite(alpha)) {
    if (lead) {
      st->done = 2;
      st->halt = 1;
    }
    return;
  }
  double v[4] = {0, 0, 0, 0};
  long i1_, i2_;
  xcd_chunk(n, i1_, i2_);
  for (long i = i1_ + threadIdx.x; i < i2_; i += WG) {
    const double xi = x[i] + alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    x[i] = xi;
    r[i] = ri;
    v[0] += dinv[i] * ri * ri;
    v[1] += ri * ri;
    v[2] += xi * xi;
  }
  block_partials<4>(v, PART_B2(part, (k + 1) & 1));
  if (lead) {
    st->pqs[slot] = pq;
    st->alpha = alpha;
    st->iter = k + 1;
  }
}

// --------------------------------------------------------------------------
// FCG(1) in three kernels per iteration (one kept direction).  With
// q_{k-1} = A p_{k-1} stored and w = A z_k formed from z_k directly,
//   beta  = -(z, q_{k-1}) / (p_{k-1}, q_{k-1}),
//   p_k   = z + beta p_{k-1},   q_k = w + beta q_{k-1}  (= A p_k),
//   (p_k, r_k) = (z, r) + beta (p_{k-1}, r_k),
//   (p_k, q_k) = (z, w) + 2 beta (z, q_{k-1}) + beta^2 (p_{k-1}, q_{k-1})
// (A symmetric: (p_{k-1}, A z) = (A p_{k-1}, z)), so alpha is known when the
// update starts, and one vector pass does p, q, x, r, the next network input
// and the partials the next iteration needs.  The update's (p, r_{k+1}) and
// (p, q) are summed directly, not carried.  Partial regions ([value][grid]
// with the writer's grid): head H at 0, 6 values -- (r, dinv r), (r, r),
// (x, x), (b, b), (p_{k-1}, r_k), (p_{k-1}, q_{k-1}) -- from the start, the
// update or a refresh (grid in st->hgrid); Z in region 6, (z, r) and
// (z, q_{k-1}) from the combine; W in region 7, (z, w) from the SpMV.
// --------------------------------------------------------------------------
#define F1_H(p) (p)
#define F1_UP 2   /* entry pairs per thread and pass of the update */
#define F1_Z(p) ((p) + 6L * 65536)
#define F1_W(p) ((p) + 7L * 65536)

// network input scale: rms(sqrt(dinv) r) of the iteration before (the
// network is linear; the scale only keeps its f32 activations near 1)
__device__ __forceinline__ double f1_scale(double rdr, long n)
{
  const double rms = sqrt(rdr / (double)n);
  return (rms > 0.0 && isfinite(rms)) ? rms : 1.0;
}

// channel c of node v of the network input, split format [V][3 parts][4]:
// the exact truncation split of f (hi + mid + lo = f)
__device__ __forceinline__ void f1_put_input(unsigned short *x0, long i, float f)
{
  const long v = i / 3, c = i - 3 * v;
  const unsigned u = __float_as_uint(f), h = u & 0xffff0000u;
  const float r1 = f - __uint_as_float(h);
  const unsigned m = __float_as_uint(r1) & 0xffff0000u;
  const float r2 = r1 - __uint_as_float(m);
  unsigned short *o = x0 + v * 12 + c;
  o[0] = (unsigned short)(h >> 16);
  o[4] = (unsigned short)(m >> 16);
  o[8] = (unsigned short)(__float_as_uint(r2) >> 16);
}

// the network input of DOF i from sdr = sqrt(dinv) r: split f32 scaled by
// 1/s (the matrix-core path), or FP64, unscaled and masked (the FP64 path)
__device__ __forceinline__ void f1_input(unsigned short *x0, double *v64, const float *mask, long i, double sdr,
                                         double inv)
{
  if (v64) v64[i] = (double)mask[i] * sdr;
  else f1_put_input(x0, i, (float)(sdr * inv));
}

template <int BS>
__global__ __launch_bounds__(WG) void k_f1_start(SellView M, const double *__restrict__ vals,
                                                 const double *__restrict__ b, const double *__restrict__ x,
                                                 double *r, const double *__restrict__ dinv, ugk_state *st,
                                                 double *part)
{
  double v[6] = {0, 0, 0, 0, 0, 0};
  double acc[BS];
  long row;
  if (sell_row_split<BS, double, double>(M, vals, x, acc, row)) {
#pragma unroll
    for (int k = 0; k < BS; k++) {
      const long i = row * BS + k;
      const double ri = b[i] - acc[k];
      r[i] = ri;
      v[0] += dinv[i] * ri * ri;
      v[1] += ri * ri;
      v[2] += x[i] * x[i];
      v[3] += b[i] * b[i];
    }
  }
  block_partials<6>(v, F1_H(part));
  if (blockIdx.x == 0 && threadIdx.x == 0) st->hgrid = gridDim.x;
}

// the first network input, from r_0 and the start's partials
__global__ __launch_bounds__(WG) void k_f1_input0(long n, const double *__restrict__ r,
                                                  const double *__restrict__ dinv, unsigned short *x0,
                                                  double *v64, const float *__restrict__ mask, ugk_state *st,
                                                  const double *part)
{
  double t[1];
  reduce_prev<1>(F1_H(part), (unsigned)st->hgrid, t);
  const double s = f1_scale(t[0], n), inv = 1.0 / s;
  long i1, i2;
  xcd_chunk(n, i1, i2);
  for (long i = i1 + threadIdx.x; i < i2; i += WG) f1_input(x0, v64, mask, i, sqrt(dinv[i]) * r[i], inv);
  if (blockIdx.x == 0 && threadIdx.x == 0) st->net_s = s;
}

// mixed form: r = b - A x in FP64 for the iteration the combine asked for,
// with the head partials (and the network input) taken from it; the kept
// direction's (p, r) and (p, q) are summed anew
template <int BS>
__global__ __launch_bounds__(WG) void k_f1_refresh(SellView M, const double *__restrict__ vals,
                                                   const double *__restrict__ b, const double *__restrict__ x,
                                                   double *r, const double *__restrict__ dinv,
                                                   const double *__restrict__ p, const double *__restrict__ q,
                                                   unsigned short *x0, double *v64, const float *__restrict__ mask,
                                                   ugk_state *st, double *part)
{
  const long k = st->iter;
  if (st->done || st->ru_req != k) return;   // the same for every block (the lead writes neither)
  const bool in = (x0 || v64) && !st->fb;
  const double inv = 1.0 / st->net_s;
  double v[6] = {0, 0, 0, 0, 0, 0};
  double acc[BS];
  long row;
  if (sell_row_split<BS, double, double>(M, vals, x, acc, row)) {
#pragma unroll
    for (int c = 0; c < BS; c++) {
      const long i = row * BS + c;
      const double ri = b[i] - acc[c];
      r[i] = ri;
      v[0] += dinv[i] * ri * ri;
      v[1] += ri * ri;
      v[2] += x[i] * x[i];
      v[4] += p[i] * ri;
      v[5] += p[i] * q[i];
      if (in) f1_input(x0, v64, mask, i, sqrt(dinv[i]) * ri, inv);
    }
  }
  block_partials<6>(v, F1_H(part));
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->ru_at = k;
    st->n_ru += 1;
    st->hgrid = gridDim.x;
    if (st->restart) st->restart_at = k;
  }
}

__global__ __launch_bounds__(WG) void k_f1_combine(long n, const double *__restrict__ r,
                                                   const double *__restrict__ dinv,
                                                   const float *__restrict__ mask,
                                                   const float *__restrict__ out,
                                                   const double *__restrict__ zc,
                                                   const double *__restrict__ q, double *__restrict__ z,
                                                   ugk_state *st, double *part)
{
  if (st->done) return;
  const long k = st->iter;
  double t[6];
  reduce_prev<6>(F1_H(part), (unsigned)st->hgrid, t);
  // mixed form (max_gap > 0): convergence is only accepted on a residual
  // recomputed in FP64 (the refresh of this iteration, or iteration 0)
  const bool mixed = st->max_gap > 0;
  const bool fresh = k == 0 || (mixed && st->ru_at == k);
  const double bb = (k == 0) ? t[3] : st->bb;
  int done = pcg_test(st, k, t[1], t[2], bb, t[0]);
  long req = -1;
  if (mixed && !fresh) {
    if (done == 1) {
      done = 0;
      req = k + 1;
    } else if (done == 0 && (t[1] < st->delta * st->delta * st->rr_true || k - st->ru_at >= st->max_gap))
      req = k + 1;
  }
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (lead) {
    st->cur = k;
    st->rr = t[1];
    st->xx = t[2];
    if (k == 0) st->bb = bb;
    if (mixed && fresh) st->rr_true = t[1];
    st->rdr = t[0];
    st->pr_old = t[4];
    st->pq_old = t[5];
    st->zgrid = gridDim.x;
    if (req >= 0) st->ru_req = req;
    if (done) {
      st->done = done;
      st->halt = 1;
    }
  }
  if (done) return;
  const bool net = out && !st->fb;            // the network ran for this iteration
  const bool useq = !(k == 0 || k == st->restart_at);
  const double sc = st->net_s;
  double v[2] = {0, 0};
  long i1, i2;
  xcd_chunk(n, i1, i2);
  for (long i = i1 + threadIdx.x; i < i2; i += WG) {
    const double ri = r[i], di = dinv[i];
    double zi = di * ri;
    if (net) {
      const long vx = i / 3;
      zi += sc * (double)mask[i] * sqrt(di) * (double)out[vx * 4 + (i - 3 * vx)];
    }
    if (zc) zi += zc[i];
    z[i] = zi;
    v[0] += zi * ri;
    if (useq) v[1] += zi * q[i];
  }
  block_partials<2>(v, F1_Z(part));
}

// The same combine with the network's last layer in it: y = C0^T t0, the
// stride-2 transpose from coarse level 0 (CIN = 16 or 32 channels) to the
// node grid, 3 outputs, on the vector ALUs in the activations' type TA (f32,
// the input scaled by st->net_s; or FP64, unscaled).  A node f = 2 c + p of parity class
// p takes, per axis, the coarse node c (offset 0) when p = 0, and c (offset
// +1) and c + 1 (offset -1, when it exists) when p = 1: 1 to 8 taps.  A
// block of 8 waves takes C0T_NODES coarse nodes, wave w the nodes of class w, in
// rounds of 16 (so a wave's taps are compile-time and its lanes never
// diverge); four lanes share a node, each with a quarter of the channels, all
// its gathers and its epilogue loads in flight together; the four partial
// sums meet by lane shuffles in a fixed order, and lane q < 3 forms z for
// DOF 3 f + q.  The weights ([27][cop][cip] bf16, the adjoint's layout) are
// widened to TA in LDS as [27][3][CIN].  Replaces the last network stage
// and k_f1_combine.
#define C0T_CIN 32          /* the largest CIN */
#define C0T_TB 512
#define C0T_NODES 128     /* coarse nodes per block (a multiple of 16): fewer blocks
                             pay the start (partials, weights) fewer times; 64:
                             +1.5 us per iteration at 1M DOF, 16: +24 us; 256
                             (FP64 t0): +11 us */

template <int PZ, int PY, int PX, int CIN>
__device__ __forceinline__ void c0t_class(const float *__restrict__ x, int ldx, const float *sw, ugk_grid3 gc, int cz,
                                          int cy, int cx, bool fok, int qd, float (&y)[3])
{
  constexpr int NZ = PZ + 1, NY = PY + 1, NX = PX + 1, NT = NZ * NY * NX, CQ = CIN / 4;
  float4 xa[NT], xb[NT];
  bool ok[NT];
#pragma unroll
  for (int j = 0; j < NT; j++) {
    const int iz = j / (NY * NX), iy = (j / NX) % NY, ix = j % NX;
    const int sz = cz + iz, sy = cy + iy, sx = cx + ix;   // offset -1 reads c + 1
    ok[j] = fok && sz < gc.D && sy < gc.H && sx < gc.W;
    xa[j] = make_float4(0.f, 0.f, 0.f, 0.f);
    xb[j] = xa[j];
    if (ok[j]) {
      const float *xs = x + (((long)sz * gc.H + sy) * gc.W + sx) * ldx + CQ * qd;
      xa[j] = *(const float4 *)xs;
      if constexpr (CQ == 8) xb[j] = *(const float4 *)(xs + 4);
    }
  }
#pragma unroll
  for (int j = 0; j < NT; j++) {
    const int iz = j / (NY * NX), iy = (j / NX) % NY, ix = j % NX;
    const int tz = PZ ? 2 - 2 * iz : 1, ty = PY ? 2 - 2 * iy : 1, tx = PX ? 2 - 2 * ix : 1;
    const float *w = sw + (9 * tz + 3 * ty + tx) * 3 * CIN + CQ * qd;
#pragma unroll
    for (int o = 0; o < 3; o++) {
      const float4 w0 = *(const float4 *)(w + o * CIN);
      float acc = w0.x * xa[j].x + w0.y * xa[j].y + w0.z * xa[j].z + w0.w * xa[j].w;
      if constexpr (CQ == 8) {
        const float4 w1 = *(const float4 *)(w + o * CIN + 4);
        acc += w1.x * xb[j].x + w1.y * xb[j].y + w1.z * xb[j].z + w1.w * xb[j].w;
      }
      y[o] += acc;
    }
  }
}

// the same for FP64 activations (the FP64 inference path): CQ doubles per lane
template <int PZ, int PY, int PX, int CIN>
__device__ __forceinline__ void c0t_class64(const double *__restrict__ x, int ldx, const double *sw, ugk_grid3 gc,
                                            int cz, int cy, int cx, bool fok, int qd, double (&y)[3])
{
  constexpr int NZ = PZ + 1, NY = PY + 1, NX = PX + 1, NT = NZ * NY * NX, CQ = CIN / 4, CH = CQ / 2;
  dbl2 xa[NT][CH];
#pragma unroll
  for (int j = 0; j < NT; j++) {
    const int iz = j / (NY * NX), iy = (j / NX) % NY, ix = j % NX;
    const int sz = cz + iz, sy = cy + iy, sx = cx + ix;   // offset -1 reads c + 1
    const bool ok = fok && sz < gc.D && sy < gc.H && sx < gc.W;
#pragma unroll
    for (int k = 0; k < CH; k++) xa[j][k] = dbl2{0.0, 0.0};
    if (ok) {
      const double *xs = x + (((long)sz * gc.H + sy) * gc.W + sx) * ldx + CQ * qd;
#pragma unroll
      for (int k = 0; k < CH; k++) xa[j][k] = *(const dbl2 *)(xs + 2 * k);
    }
  }
#pragma unroll
  for (int j = 0; j < NT; j++) {
    const int iz = j / (NY * NX), iy = (j / NX) % NY, ix = j % NX;
    const int tz = PZ ? 2 - 2 * iz : 1, ty = PY ? 2 - 2 * iy : 1, tx = PX ? 2 - 2 * ix : 1;
    const double *w = sw + (9 * tz + 3 * ty + tx) * 3 * CIN + CQ * qd;
#pragma unroll
    for (int o = 0; o < 3; o++) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < CH; k++) acc += w[o * CIN + 2 * k] * xa[j][k].x + w[o * CIN + 2 * k + 1] * xa[j][k].y;
      y[o] += acc;
    }
  }
}

template <int CIN, typename TA>
__global__ __launch_bounds__(C0T_TB) void k_f1_combine_c0t(long n, const double *__restrict__ r,
                                                           const double *__restrict__ dinv,
                                                           const float *__restrict__ mask, ugk_conv_desc cd,
                                                           const double *__restrict__ zc,
                                                           const double *__restrict__ q, double *__restrict__ z,
                                                           ugk_state *st, double *part, int npb)
{
  __shared__ TA sw[27 * 3 * C0T_CIN];
  if (st->done) return;
  const long k = st->iter;
  double t[6];
  reduce_prev_tb<6, C0T_TB>(F1_H(part), (unsigned)st->hgrid, t);
  const bool mixed = st->max_gap > 0;
  const bool fresh = k == 0 || (mixed && st->ru_at == k);
  const double bb = (k == 0) ? t[3] : st->bb;
  int done = pcg_test(st, k, t[1], t[2], bb, t[0]);
  long req = -1;
  if (mixed && !fresh) {
    if (done == 1) {
      done = 0;
      req = k + 1;
    } else if (done == 0 && (t[1] < st->delta * st->delta * st->rr_true || k - st->ru_at >= st->max_gap))
      req = k + 1;
  }
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (lead) {
    st->cur = k;
    st->rr = t[1];
    st->xx = t[2];
    if (k == 0) st->bb = bb;
    if (mixed && fresh) st->rr_true = t[1];
    st->rdr = t[0];
    st->pr_old = t[4];
    st->pq_old = t[5];
    st->zgrid = gridDim.x;
    if (req >= 0) st->ru_req = req;
    if (done) {
      st->done = done;
      st->halt = 1;
    }
  }
  if (done) return;
  const bool net = !st->fb;                   // the network ran for this iteration
  if (net) {
    const unsigned short *w = cd.w;
    for (int i = threadIdx.x; i < 27 * 3 * CIN; i += C0T_TB) {
      const int tp = i / (3 * CIN), o = (i / CIN) % 3, c = i % CIN;
      sw[i] = (TA)__uint_as_float((unsigned)w[((long)tp * cd.cop + o) * cd.cip + c] << 16);
    }
  }
  __syncthreads();
  const bool useq = !(k == 0 || k == st->restart_at);
  const double sc = sizeof(TA) == 8 ? 1.0 : st->net_s;   // FP64 input: unscaled
  const TA *x = (const TA *)cd.x;
  const ugk_grid3 gc = cd.gi, gf = cd.go;
  const int lane = threadIdx.x & 63, cls = threadIdx.x >> 6, qd = lane & 3;
  const int pz = cls >> 2, py = (cls >> 1) & 1, px = cls & 1;
  const unsigned Vc = (unsigned)gc.D * gc.H * gc.W, HWc = (unsigned)gc.H * gc.W, Wc = (unsigned)gc.W;
  double v[2] = {0, 0};
#pragma unroll 1   // one copy of the eight class bodies: the code of all classes is live on every CU
  for (int rd = 0; rd < npb / 16; rd++) {
    const unsigned c = (unsigned)blockIdx.x * npb + rd * 16 + (lane >> 2);
    const unsigned cz = c / HWc, cr = c - cz * HWc, cy = cr / Wc, cx = cr - cy * Wc;
    const int fz = 2 * cz + pz, fy = 2 * cy + py, fx = 2 * cx + px;
    const bool fok = c < Vc && fz < gf.D && fy < gf.H && fx < gf.W;
    const long i = fok ? 3 * (((long)fz * gf.H + fy) * gf.W + fx) + (qd < 3 ? qd : 0) : 0;
    const bool mine = fok && qd < 3;
    double ri = 0.0, di = 0.0, qi = 0.0, zci = 0.0;
    float mi = 0.f;
    if (mine) {          // the epilogue's operands, in flight with the gathers
      ri = r[i];
      di = dinv[i];
      mi = mask[i];
      if (useq) qi = q[i];
      if (zc) zci = zc[i];
    }
    TA y[3] = {0, 0, 0};
    if (net) {
      if constexpr (sizeof(TA) == 8) {
        switch (cls) {                        // wave-uniform
          case 0: c0t_class64<0, 0, 0, CIN>(x, cd.ldx, sw, gc, cz, cy, cx, fok, qd, y); break;
          case 1: c0t_class64<0, 0, 1, CIN>(x, cd.ldx, sw, gc, cz, cy, cx, fok, qd, y); break;
          case 2: c0t_class64<0, 1, 0, CIN>(x, cd.ldx, sw, gc, cz, cy, cx, fok, qd, y); break;
          case 3: c0t_class64<0, 1, 1, CIN>(x, cd.ldx, sw, gc, cz, cy, cx, fok, qd, y); break;
          case 4: c0t_class64<1, 0, 0, CIN>(x, cd.ldx, sw, gc, cz, cy, cx, fok, qd, y); break;
          case 5: c0t_class64<1, 0, 1, CIN>(x, cd.ldx, sw, gc, cz, cy, cx, fok, qd, y); break;
          case 6: c0t_class64<1, 1, 0, CIN>(x, cd.ldx, sw, gc, cz, cy, cx, fok, qd, y); break;
          default: c0t_class64<1, 1, 1, CIN>(x, cd.ldx, sw, gc, cz, cy, cx, fok, qd, y); break;
        }
      } else {
        switch (cls) {                        // wave-uniform
          case 0: c0t_class<0, 0, 0, CIN>(x, cd.ldx, sw, gc, cz, cy, cx, fok, qd, y); break;
          case 1: c0t_class<0, 0, 1, CIN>(x, cd.ldx, sw, gc, cz, cy, cx, fok, qd, y); break;
          case 2: c0t_class<0, 1, 0, CIN>(x, cd.ldx, sw, gc, cz, cy, cx, fok, qd, y); break;
          case 3: c0t_class<0, 1, 1, CIN>(x, cd.ldx, sw, gc, cz, cy, cx, fok, qd, y); break;
          case 4: c0t_class<1, 0, 0, CIN>(x, cd.ldx, sw, gc, cz, cy, cx, fok, qd, y); break;
          case 5: c0t_class<1, 0, 1, CIN>(x, cd.ldx, sw, gc, cz, cy, cx, fok, qd, y); break;
          case 6: c0t_class<1, 1, 0, CIN>(x, cd.ldx, sw, gc, cz, cy, cx, fok, qd, y); break;
          default: c0t_class<1, 1, 1, CIN>(x, cd.ldx, sw, gc, cz, cy, cx, fok, qd, y); break;
        }
      }
#pragma unroll
      for (int o = 0; o < 3; o++) {
        y[o] += __shfl_xor(y[o], 1);
        y[o] += __shfl_xor(y[o], 2);
      }
    }
    if (mine) {
      double zi = di * ri;
      if (net) zi += sc * (double)mi * sqrt(di) * (double)(qd == 0 ? y[0] : (qd == 1 ? y[1] : y[2]));
      zi += zci;
      z[i] = zi;
      v[0] += zi * ri;
      v[1] += zi * qi;
    }
  }
  block_partials_tb<2, C0T_TB>(v, F1_Z(part));
}

// w = A z with (z, w); T = float: the FP32 copy of A, FP64 vectors and sums
template <int BS, typename T>
__global__ __launch_bounds__(WG) void k_f1_w(SellView M, const T *__restrict__ vals,
                                             const double *__restrict__ z, double *__restrict__ w,
                                             const ugk_state *st, double *part)
{
  if (st->done) return;
  double v[1] = {0.0};
  double acc[BS];
  long row;
  if (sell_row_split<BS, T, double>(M, vals, z, acc, row)) {
#pragma unroll
    for (int c = 0; c < BS; c++) {
      const long i = row * BS + c;
      w[i] = acc[c];
      v[0] += z[i] * acc[c];
    }
  }
  block_partials<1>(v, F1_W(part));
}

__global__ __launch_bounds__(WG) void k_f1_update(long n, double *__restrict__ x, double *__restrict__ r,
                                                  double *__restrict__ p, double *__restrict__ q,
                                                  const double *__restrict__ z, const double *__restrict__ w,
                                                  const double *__restrict__ dinv, unsigned short *x0,
                                                  double *v64, const float *__restrict__ mask,
                                                  ugk_state *st, double *part, unsigned GW)
{
  if (st->done) return;
  const long k = st->cur;
  double tz[2], tw[1];
  reduce_prev<2>(F1_Z(part), (unsigned)st->zgrid, tz);
  reduce_prev<1>(F1_W(part), GW, tw);
  const bool restart = k == 0 || k == st->restart_at;
  const double zr = tz[0], zq = tz[1], zw = tw[0];
  const double pq_old = st->pq_old;
  const double beta = restart ? 0.0 : -zq / pq_old;
  const double pr = restart ? zr : zr + beta * st->pr_old;
  const double pq = restart ? zw : zw + 2.0 * beta * zq + beta * beta * pq_old;
  const bool ok = isfinite(beta) && isfinite(pr) && isfinite(pq) && pr > 0.0 && pq > 0.0;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (!ok) {
    // not a descent direction: iteration k moves nothing (x, r and the head
    // partials stay), and iteration k + 1 starts afresh with z = dinv r, the
    // network off from here on; with z = dinv r already, a breakdown.  The
    // lead writes only fields this kernel's blocks do not read (restart_at
    // becomes k + 1, never k), so every block takes this same branch.
    if (lead) {
      if (!st->fb) {
        st->fb = 1;
        st->fb_iter = k;
        st->restart_at = k + 1;
        st->iter = k + 1;
      } else {
        st->done = 2;
      }
      st->halt = 1;
    }
    return;
  }
  const double alpha = pr / pq;
  const bool in = (x0 || v64) && !st->fb;
  const double s = f1_scale(st->rdr, n), inv = 1.0 / s;
  double v[6] = {0, 0, 0, 0, 0, 0};
  long i1, i2;
  xcd_chunk_even(n, i1, i2);
  // two entries per lane (16-byte accesses), F1_UP pairs per thread and
  // pass with all loads issued before the first use
  long i = i1 + 2 * threadIdx.x;
  for (; i + 2L * WG * (F1_UP - 1) + 1 < i2; i += 2L * WG * F1_UP) {
    dbl2 zz[F1_UP], pp[F1_UP], ww[F1_UP], qq[F1_UP], xx[F1_UP], rr[F1_UP], dd[F1_UP];
#pragma unroll
    for (int u = 0; u < F1_UP; u++) {
      const long j = i + 2L * WG * u;
      zz[u] = *(const dbl2 *)(z + j);
      ww[u] = *(const dbl2 *)(w + j);
      pp[u] = restart ? dbl2{0.0, 0.0} : *(const dbl2 *)(p + j);
      qq[u] = restart ? dbl2{0.0, 0.0} : *(const dbl2 *)(q + j);
      xx[u] = *(const dbl2 *)(x + j);
      rr[u] = *(const dbl2 *)(r + j);
      dd[u] = *(const dbl2 *)(dinv + j);
    }
#pragma unroll
    for (int u = 0; u < F1_UP; u++) {
      const long j = i + 2L * WG * u;
      const dbl2 pi = zz[u] + beta * pp[u], qi = ww[u] + beta * qq[u];
      const dbl2 xi = xx[u] + alpha * pi, ri = rr[u] - alpha * qi, di = dd[u];
      *(dbl2 *)(p + j) = pi;
      *(dbl2 *)(q + j) = qi;
      *(dbl2 *)(x + j) = xi;
      *(dbl2 *)(r + j) = ri;
      v[0] += di.x * ri.x * ri.x + di.y * ri.y * ri.y;
      v[1] += ri.x * ri.x + ri.y * ri.y;
      v[2] += xi.x * xi.x + xi.y * xi.y;
      v[4] += pi.x * ri.x + pi.y * ri.y;
      v[5] += pi.x * qi.x + pi.y * qi.y;
      if (in) {
        if (v64) {
          const float2 m2 = *(const float2 *)(mask + j);
          *(dbl2 *)(v64 + j) = dbl2{(double)m2.x * sqrt(di.x) * ri.x, (double)m2.y * sqrt(di.y) * ri.y};
        } else {
          f1_put_input(x0, j, (float)(sqrt(di.x) * ri.x * inv));
          f1_put_input(x0, j + 1, (float)(sqrt(di.y) * ri.y * inv));
        }
      }
    }
  }
  for (; i < i2; i += 2L * WG) {
    const int m = i + 1 < i2 ? 2 : 1;
    for (int c = 0; c < m; c++) {
      const long j = i + c;
      const double pi = restart ? z[j] : z[j] + beta * p[j];
      const double qi = restart ? w[j] : w[j] + beta * q[j];
      const double xi = x[j] + alpha * pi;
      const double ri = r[j] - alpha * qi;
      const double di = dinv[j];
      p[j] = pi;
      q[j] = qi;
      x[j] = xi;
      r[j] = ri;
      v[0] += di * ri * ri;
      v[1] += ri * ri;
      v[2] += xi * xi;
      v[4] += pi * ri;
      v[5] += pi * qi;
      if (in) f1_input(x0, v64, mask, j, sqrt(di) * ri, inv);
    }
  }
  block_partials<6>(v, F1_H(part));
  if (lead) {
    st->iter = k + 1;
    st->alpha = alpha;
    st->beta = beta;
    st->net_s = s;
    st->hgrid = gridDim.x;
  }
}


// --------------------------------------------------------------------------
// CG with the SSOR (omega = 1, symmetric Gauss-Seidel) preconditioner,
// femEqsCGwSSOR (fem/fem_eqs.c:999-1188), on a colour-permuted copy of A:
// block rows are ordered colour by colour (no two rows of a colour couple),
// so a sweep is one launch per colour over a contiguous range [r0, r1) of
// rows.  Forward: zz_I = D_I^-1 (r_I - sum_{J < r0} A_IJ zz_J), the 3x3
// diagonal block solved lower-triangular in place; backward, colours in
// reverse: z_I = zz_I - D_I^-1 (sum_{J >= r1} A_IJ z_J + upper part of the
// diagonal block).  Rows of later (forward) or earlier (backward) colours
// are never read, so nothing is cleared between sweeps.
// Partials: start -> test: B2(0); upd(k) -> test(k+1): B2((k+1) & 1), with
// (., r.r, x.x, b.b); dot -> p: S (r.z); q -> upd: A (p.q).
// --------------------------------------------------------------------------
template <int BS, int FWD>
__global__ __launch_bounds__(WG) void k_sgs(SellView M, const double *__restrict__ vals, long r0, long r1,
                                            const double *__restrict__ src, const double *__restrict__ dinv,
                                            double *y, const ugk_state *st)
{
  if (st->done) return;
  const int lane = threadIdx.x & 63;
  const long sl = r0 / 64 + (long)blockIdx.x * NW + (threadIdx.x >> 6);
  const long I = sl * 64 + lane;
  if (I < r0 || I >= r1) return;           // no barriers below
  const int L = M.rowlen[I];
  const long so = M.soff[sl];
  const int *cp = M.cols + so * 64 + lane;
  double acc[BS], D[BS][BS];
#pragma unroll
  for (int a = 0; a < BS; a++) {
    acc[a] = 0.0;
#pragma unroll
    for (int b = 0; b < BS; b++) D[a][b] = 0.0;
  }
  for (int k = 0; k < L; k++) {
    const long J = cp[(long)k * 64];
    if (J == I) {
#pragma unroll
      for (int q = 0; q < BS * BS; q++) D[q / BS][q % BS] = vals[vpos<BS>(so + k, q, lane)];
    } else if (FWD ? J < r0 : J >= r1) {
      double yv[BS];
#pragma unroll
      for (int c = 0; c < BS; c++) yv[c] = y[J * BS + c];
#pragma unroll
      for (int a = 0; a < BS; a++)
#pragma unroll
        for (int c = 0; c < BS; c++) acc[a] += vals[vpos<BS>(so + k, a * BS + c, lane)] * yv[c];
    }
  }
  double yl[BS];
  if (FWD) {
#pragma unroll
    for (int d = 0; d < BS; d++) {
      double t = acc[d];
#pragma unroll
      for (int e = 0; e < d; e++) t += D[d][e] * yl[e];
      yl[d] = dinv[I * BS + d] * (src[I * BS + d] - t);
      y[I * BS + d] = yl[d];
    }
  } else {
#pragma unroll
    for (int d = BS - 1; d >= 0; d--) {
      double t = acc[d];
#pragma unroll
      for (int e = d + 1; e < BS; e++) t += D[d][e] * yl[e];
      yl[d] = src[I * BS + d] - dinv[I * BS + d] * t;
      y[I * BS + d] = yl[d];
    }
  }
}

// convergence test on the previous update's partials (iteration st->iter)
__global__ __launch_bounds__(WG) void k_ssor_test(ugk_state *st, const double *part, unsigned G0, unsigned GB)
{
  if (st->done) return;
  const long k = st->iter;
  double t[4];
  reduce_prev<4>(PART_B2(part, k & 1), k == 0 ? G0 : GB, t);
  const double bb = (k == 0) ? t[3] : st->bb;
  const int done = pcg_test(st, k, t[1], t[2], bb, t[1]);
  if (threadIdx.x == 0) {
    st->cur = k;
    st->rr = t[1];
    st->xx = t[2];
    if (k == 0) st->bb = bb;
    if (done) st->done = done;
  }
}

__global__ __launch_bounds__(WG) void k_ssor_dot(long n, const double *__restrict__ r, const double *__restrict__ z,
                                                 const ugk_state *st, double *part)
{
  if (st->done) return;
  double v[1] = {0.0};
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) v[0] += r[i] * z[i];
  block_partials<1>(v, PART_S(part));
}

// beta = (r,z) / (r,z)_prev, p = z + beta p (p = z at the first iteration)
__global__ __launch_bounds__(WG) void k_ssor_p(long n, const double *__restrict__ z, double *p, ugk_state *st,
                                               const double *part, unsigned GV)
{
  if (st->done) return;
  const long k = st->cur;
  double t[1];
  reduce_prev<1>(PART_S(part), GV, t);
  const double ro = t[0];
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (!isfinite(ro)) {
    if (lead) st->done = 2;
    return;
  }
  const double beta = k == 0 ? 0.0 : ro / st->rz[(k + 1) & 1];
  if (lead) {
    st->rz[k & 1] = ro;
    st->rzB = ro;
    st->beta = beta;
  }
  long i1, i2;
  xcd_chunk(n, i1, i2);
  for (long i = i1 + threadIdx.x; i < i2; i += WG)
    p[i] = k == 0 ? z[i] : z[i] + beta * p[i];
}

// x += alpha p, r -= alpha q with the partials (r dinv r, r r, x x) of the
// next test; dinv may be NULL (SSOR, two-level: the first partial stays 0)
// Jacobi direction from the head's (r, dinv r): beta = ro / ro_prev,
// p = dinv r + beta p (p = dinv r at the first iteration); z is never stored
__global__ __launch_bounds__(WG) void k_jac_p(long n, const double *__restrict__ r, const double *__restrict__ dinv,
                                              double *p, ugk_state *st)
{
  if (st->done) return;
  const long k = st->cur;
  const double ro = st->aux[6];
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (!isfinite(ro)) {
    if (lead) st->done = 2;
    return;
  }
  const double beta = k == 0 ? 0.0 : ro / st->rz[(k + 1) & 1];
  if (lead) {
    st->rz[k & 1] = ro;
    st->rzB = ro;
    st->beta = beta;
  }
  long i1, i2;
  xcd_chunk(n, i1, i2);
  for (long i = i1 + threadIdx.x; i < i2; i += WG) p[i] = k == 0 ? dinv[i] * r[i] : dinv[i] * r[i] + beta * p[i];
}

// Head and Jacobi direction of the FP64-vector mixed PCG in one launch:
// every block sums the partials of (r, dinv r), (r, r), (x, x), (b, b) (the
// refresh's, G0 blocks, when it ran for this iteration), tests convergence
// (accepted on a refreshed residual only), and block 0 records the refresh
// request: the recursive ||r|| has fallen by delta since the last FP64
// residual, max_gap iterations, or a convergence to confirm.  Then
// p = dinv r + beta p.  Fields other blocks read are not written here
// (rz by parity), so every block takes the same decisions.
__global__ __launch_bounds__(WG) void k_mx64_p(long n, const double *__restrict__ r,
                                               const double *__restrict__ dinv, double *p, ugk_state *st,
                                               const double *part, unsigned G0, unsigned GB)
{
  if (st->done) return;
  const long k = st->iter;
  const bool fresh = k == 0 || st->ru_at == k;
  double t[4];
  reduce_prev<4>(PART_B2(part, k & 1), fresh ? G0 : GB, t);
  const double bb = (k == 0) ? t[3] : st->bb;
  const double ro = t[0];
  int done = pcg_test(st, k, t[1], t[2], bb, ro);
  const bool confirm = done == 1 && !fresh;
  if (confirm) done = 0;
  if (!done && !isfinite(ro)) done = 2;
  const double beta = k == 0 ? 0.0 : ro / st->rz[(k + 1) & 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->cur = k;
    st->rr = t[1];
    st->xx = t[2];
    if (k == 0) st->bb = bb;
    if (fresh) {
      st->rr_true = t[1];
      st->ru_req = -1;   // served
    } else if (confirm || (done == 0 && (t[1] < st->delta * st->delta * st->rr_true ||
                                         k - st->ru_at >= st->max_gap)))
      st->ru_req = k + 1;
    if (done) st->done = done;
    else {
      st->rz[k & 1] = ro;
      st->rzB = ro;
      st->beta = beta;
    }
  }
  if (done) return;
  long i1, i2;
  xcd_chunk_even(n, i1, i2);   // two entries per lane, 16-byte accesses
  for (long i = i1 + 2 * threadIdx.x; i < i2; i += 2 * WG) {
    if (i + 1 < i2) {
      dbl2 pn = *(const dbl2 *)(dinv + i) * *(const dbl2 *)(r + i);
      if (k > 0) pn += beta * *(const dbl2 *)(p + i);
      *(dbl2 *)(p + i) = pn;
    } else
      p[i] = k == 0 ? dinv[i] * r[i] : dinv[i] * r[i] + beta * p[i];
  }
}

__global__ __launch_bounds__(WG) void k_ssor_upd(long n, double *__restrict__ x, double *__restrict__ r,
                                                 const double *__restrict__ p, const double *__restrict__ q,
                                                 ugk_state *st, double *part, unsigned GA,
                                                 const double *__restrict__ dinv)
{
  if (st->done) return;
  const long k = st->cur;
  double t[1];
  reduce_prev<1>(PART_A(part), GA, t);
  const double alpha = st->rzB / t[0];
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (!isfinite(alpha)) {
    if (lead) st->done = 2;
    return;
  }
  double v[4] = {0, 0, 0, 0};
  long i1, i2;
  xcd_chunk(n, i1, i2);
  for (long i = i1 + threadIdx.x; i < i2; i += WG) {
    const double xi = x[i] + alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    x[i] = xi;
    r[i] = ri;
    if (dinv) v[0] += dinv[i] * ri * ri;
    v[1] += ri * ri;
    v[2] += xi * xi;
  }
  block_partials<4>(v, PART_B2(part, (k + 1) & 1));
  if (lead) {
    st->alpha = alpha;
    st->iter = k + 1;
  }
}

// --------------------------------------------------------------------------
// CG with the SSOR preconditioner by Eisenstat's trick (colour order, scalar
// D, L, U of Kc = L + D + U, W = D + L, M = W D^-1 W^T).  CG runs on
// A^ = D^1/2 W^-1 A W^-T D^1/2 for the correction x^ of x = x0 + W^-T D^1/2 x^:
// since A = W + W^T - D, A^ p = D^1/2 (t + s) with W^T t = D^1/2 p and
// W s = D^1/2 p - D t, one backward and one forward triangular sweep per
// iteration and no separate product with A.  A sweep is one launch per
// colour over its contiguous rows; a row reads only the slots below its
// diagonal slot (forward, earlier colours) or above it (backward, later
// colours) -- dpos[I] -- so the two sweeps together read Kc once.  One
// workgroup per slice: its 4 waves take a quarter of each row's slots and
// meet in LDS; the 3x3 (or 1x1) diagonal block is solved triangularly per
// row.  The iterates are those of SSOR-PCG; the residual of the original
// system, r = W D^-1/2 r^, is formed and tested every few iterations.
//
// Partials: sweeps -> PART_S (one per block, launches back to back, pbase
// offsets); update(k) -> PART_B2((k+1)&1) (r^, r^); residual -> PART_R.
// --------------------------------------------------------------------------
enum { EIS_BWD_P = 0, EIS_FWD_Q = 1, EIS_FWD_R0 = 2, EIS_BWD_X = 3 };

template <int BS, int MODE>
__global__ __launch_bounds__(WG) void k_eis_sweep(SellView M, const double *__restrict__ vals,
                                                  const int *__restrict__ dpos, long r0, long r1, long sl0,
                                                  const double *__restrict__ in1, const double *__restrict__ in2,
                                                  double *y, double *out1, double *out2, ugk_state *st,
                                                  double *part, unsigned GB, int gate)
{
  constexpr bool FWD = (MODE == EIS_FWD_Q || MODE == EIS_FWD_R0);
  if (gate && st->done) return;            // the same for every block
  __shared__ double sh[NW - 1][BS][64];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long sl = sl0 + blockIdx.x;
  const long I = sl * 64 + lane;
  const bool act = I >= r0 && I < r1;
  // backward sweep of an iteration after the first: the direction update
  // p^ = r^ + beta p^ of this row first (beta from the update's partials)
  double beta = 0.0;
  if (MODE == EIS_BWD_P && st->iter > 0) {
    double t[1];
    reduce_prev<1>(PART_B2(part, st->iter & 1), GB, t);
    beta = t[0] / st->rzB;
  }
  double acc[BS];
#pragma unroll
  for (int r = 0; r < BS; r++) acc[r] = 0.0;
  const long so = M.soff[sl];
  int dp = 0, L = 0;
  if (act) {
    dp = dpos[I];
    L = M.rowlen[I];
  }
  const int lo = FWD ? 0 : dp + 1, hi = FWD ? dp : L, cnt = hi - lo;
  const int q0 = lo + (cnt * w) / NW, q1 = lo + (cnt * (w + 1)) / NW;
  const int *cp = M.cols + so * 64 + lane;
  for (int k = q0; k < q1; k++) {
    const long J = ldnt(cp + (long)k * 64);
    double yv[BS];
#pragma unroll
    for (int c = 0; c < BS; c++) yv[c] = y[J * BS + c];
    double v[BS * BS];
    ld_block<BS, true>(vals, so + k, lane, v);
#pragma unroll
    for (int r = 0; r < BS; r++)
#pragma unroll
      for (int c = 0; c < BS; c++) acc[r] += v[r * BS + c] * yv[c];
  }
  if (w > 0) {
#pragma unroll
    for (int r = 0; r < BS; r++) sh[w - 1][r][lane] = acc[r];
  }
  __syncthreads();
  if (w > 0) return;                       // no barrier below
  double pv = 0.0;
  if (act) {
#pragma unroll
    for (int k = 0; k < NW - 1; k++)
#pragma unroll
      for (int r = 0; r < BS; r++) acc[r] += sh[k][r][lane];
    double B[BS * BS];
    ld_block<BS, false>(vals, so + dp, lane, B);
    double dq[BS], rhs[BS], yl[BS];
#pragma unroll
    for (int d = 0; d < BS; d++) dq[d] = sqrt(B[d * BS + d]);
#pragma unroll
    for (int d = 0; d < BS; d++) {
      const long i = I * BS + d;
      if (MODE == EIS_BWD_P) {
        double ph = out1[i];
        if (st->iter > 0) {
          ph = in1[i] + beta * ph;          // in1 = r^, out1 = p^
          out1[i] = ph;
        }
        rhs[d] = dq[d] * ph;
      } else if (MODE == EIS_FWD_Q) {
        rhs[d] = dq[d] * in1[i] - B[d * BS + d] * in2[i];   // in1 = p^, in2 = t
      } else if (MODE == EIS_FWD_R0) {
        rhs[d] = in1[i];                    // r0 = b - A x0
      } else {
        rhs[d] = dq[d] * in1[i];            // in1 = x^
      }
    }
    if (FWD) {
#pragma unroll
      for (int d = 0; d < BS; d++) {
        double t = acc[d];
#pragma unroll
        for (int e = 0; e < d; e++) t += B[d * BS + e] * yl[e];
        yl[d] = (rhs[d] - t) / B[d * BS + d];
      }
    } else {
#pragma unroll
      for (int d = BS - 1; d >= 0; d--) {
        double t = acc[d];
#pragma unroll
        for (int e = d + 1; e < BS; e++) t += B[d * BS + e] * yl[e];
        yl[d] = (rhs[d] - t) / B[d * BS + d];
      }
    }
#pragma unroll
    for (int d = 0; d < BS; d++) {
      const long i = I * BS + d;
      y[i] = yl[d];
      if (MODE == EIS_FWD_Q) {
        const double qh = dq[d] * (in2[i] + yl[d]);
        out1[i] = qh;
        pv += in1[i] * qh;                  // (p^, q^)
      } else if (MODE == EIS_FWD_R0) {
        const double rh = dq[d] * yl[d];
        out1[i] = rh;
        out2[i] = rh;
        pv += rh * rh;                      // (r^, r^)
      } else if (MODE == EIS_BWD_X) {
        const double xi = in2[i] + yl[d];   // in2 = x0
        if (out1) out1[i] = xi;
        pv += xi * xi;                      // (x, x)
      }
    }
  }
  if (MODE != EIS_BWD_P) {
    pv = wave_sum(pv);
    if (lane == 0) part[blockIdx.x] = pv;
  }
}

// one block: the sum of G partials into dst[0] (a single partial for the next reader)
__global__ __launch_bounds__(WG) void k_eis_sum(const double *__restrict__ src, unsigned G, double *dst)
{
  double t[1];
  reduce_prev<1>(src, G, t);
  if (threadIdx.x == 0) dst[0] = t[0];
}

// alpha = (r^, r^) / (p^, q^); x^ += alpha p^, r^ -= alpha q^; (r^, r^) partials
__global__ __launch_bounds__(WG) void k_eis_upd(long n, double *__restrict__ xh, double *__restrict__ rh,
                                                const double *__restrict__ ph, const double *__restrict__ qh,
                                                ugk_state *st, double *part, unsigned GS, unsigned GB,
                                                int check)
{
  if (st->done) return;
  const long k = st->iter;
  double t[1], u[1];
  reduce_prev<1>(PART_S(part), GS, t);
  reduce_prev<1>(PART_B2(part, k & 1), k == 0 ? 1u : GB, u);
  const double rho = u[0], alpha = rho / t[0];
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (!(t[0] > 0.0) || !isfinite(alpha)) {
    if (lead) st->done = 2;
    return;
  }
  double v[1] = {0.0};
  long i1, i2;
  xcd_chunk(n, i1, i2);
  for (long i = i1 + threadIdx.x; i < i2; i += WG) {
    xh[i] += alpha * ph[i];
    const double ri = rh[i] - alpha * qh[i];
    rh[i] = ri;
    v[0] += ri * ri;
  }
  block_partials<1>(v, PART_B2(part, (k + 1) & 1));
  if (lead) {
    st->rzB = rho;
    st->alpha = alpha;
    st->iter = k + 1;
    if (!check && k + 1 >= st->maxit) st->done = 3;
  }
}

// r = W D^-1/2 r^ (the original residual, colour order) and its partials:
// r_I = sum of the slots below the diagonal A_IJ v_J + the diagonal block's
// lower triangle with its diagonal times v_I, v = r^ / sqrt(diag)
template <int BS>
__global__ __launch_bounds__(WG) void k_eis_resid(SellView M, const double *__restrict__ vals,
                                                  const int *__restrict__ dpos, const double *__restrict__ rh,
                                                  const double *__restrict__ dsq, ugk_state *st, double *part)
{
  if (st->done) return;
  double v[1] = {0.0};
  const int lane = threadIdx.x & 63;
  for (long sl = (long)blockIdx.x * NW + (threadIdx.x >> 6); sl < M.n_slices; sl += (long)gridDim.x * NW) {
    const long I = sl * 64 + lane;
    if (I >= M.n_brows) continue;
    const long so = M.soff[sl];
    const int dp = dpos[I];
    double acc[BS];
#pragma unroll
    for (int r = 0; r < BS; r++) acc[r] = 0.0;
    for (int k = 0; k < dp; k++) {
      const long J = M.cols[(so + k) * 64 + lane];
      double yv[BS];
#pragma unroll
      for (int c = 0; c < BS; c++) yv[c] = rh[J * BS + c] / dsq[J * BS + c];
      double a[BS * BS];
      ld_block<BS, true>(vals, so + k, lane, a);
#pragma unroll
      for (int r = 0; r < BS; r++)
#pragma unroll
        for (int c = 0; c < BS; c++) acc[r] += a[r * BS + c] * yv[c];
    }
    double B[BS * BS], yo[BS];
    ld_block<BS, false>(vals, so + dp, lane, B);
#pragma unroll
    for (int c = 0; c < BS; c++) yo[c] = rh[I * BS + c] / dsq[I * BS + c];
#pragma unroll
    for (int r = 0; r < BS; r++) {
      double t = acc[r];
#pragma unroll
      for (int c = 0; c <= r; c++) t += B[r * BS + c] * yo[c];
      v[0] += t * t;
    }
  }
  block_partials<1>(v, PART_R(part));
}

// sqrt of the scalar diagonal, colour order
template <int BS>
__global__ void k_eis_dsq(SellView M, const double *__restrict__ vals, const int *__restrict__ dpos,
                          double *dsq, ugk_state *st)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M.n_brows * BS) return;
  const long I = i / BS, d = i % BS;
  const double a = vals[vpos<BS>(M.soff[I >> 6] + dpos[I], (int)(d * BS + d), (int)(I & 63))];
  if (!(a > 0.0)) {
    st->err = 4;
    dsq[i] = 1.0;
    return;
  }
  dsq[i] = sqrt(a);
}

// convergence test on the residual's partials (and, for the reference rule,
// the (x, x) partials of the sweep that formed x)
__global__ __launch_bounds__(WG) void k_eis_test(ugk_state *st, const double *part, unsigned GR, unsigned GX)
{
  if (st->done) return;
  double t[1], u[1] = {0.0};
  reduce_prev<1>(PART_R(part), GR, t);
  if (GX) reduce_prev<1>(PART_S(part), GX, u);
  const long k = st->iter;
  const int done = pcg_test(st, k, t[0], u[0], st->bb, t[0]);
  if (threadIdx.x == 0) {
    st->rr = t[0];
    st->xx = u[0];
    if (done) st->done = done;
  }
}

// values of the colour-permuted matrix Kc from K: new block row I' is old
// row perm[I'], its column J' old column perm[J'] (found by binary search)
template <int BS>
__global__ __launch_bounds__(WG) void k_ssor_permute_vals(SellView M, const double *__restrict__ vals, SellView Mc,
                                                          double *valc, const int *__restrict__ perm)
{
  const long I = (long)blockIdx.x * WG + threadIdx.x;
  if (I >= Mc.n_brows) return;
  const long sl = I >> 6, ln = I & 63;
  const long Io = perm[I];
  const long slo = Io >> 6, lno = Io & 63;
  const int L = Mc.rowlen[I];
  for (int k = 0; k < L; k++) {
    const long at = (Mc.soff[sl] + k) * 64 + ln;
    const int Jo = perm[Mc.cols[at]];
    const int ko = find_slot(M, Io, Jo);
#pragma unroll
    for (int q = 0; q < BS * BS; q++)
      valc[vpos<BS>(Mc.soff[sl] + k, q, (int)ln)] = ko >= 0 ? vals[vpos<BS>(M.soff[slo] + ko, q, (int)lno)] : 0.0;
  }
}

// dir 0: dst[I'] = src[perm[I']] (into colour order); dir 1: dst[perm[I']] = src[I']
__global__ __launch_bounds__(WG) void k_ssor_permute_vec(long nb, int bs, const int *__restrict__ perm,
                                                         const double *__restrict__ src, double *__restrict__ dst,
                                                         int dir)
{
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < nb * bs; i += (long)gridDim.x * WG) {
    const long I = i / bs, c = i - I * bs, Io = perm[I];
    if (dir == 0) dst[i] = src[Io * bs + c];
    else dst[Io * bs + c] = src[i];
  }
}

static SellView view(const ugk_mat *K)
{
  SellView M;
  M.slen = K->d_slen;
  M.soff = K->d_soff;
  M.rowlen = K->d_rowlen;
  M.cols = K->d_cols;
  M.n_brows = K->n_brows;
  M.n_slices = K->n_slices;
  M.corr = K->bs == 3 ? K->d_corr32 : nullptr;
  M.rem = K->d_rem16;
  return M;
}
// --------------------------------------------------------------------------
// Two-level correction on a box node grid (see ugk.h): z_c = fm P A^-1 P^T
// fm r with P trilinear from a coarse node grid, A the Galerkin product
// assembled per cell from the elements' Lame constants and inverted once
// per solve by block Gauss-Jordan.  Fine node (i, j, k) belongs to the cell
// min(i / cx, Wc - 2), ... (the last cell also owns the last node plane);
// its weight for the cell's corner with offset 1 along x is t / cx with
// t = i - cx * cell, 1 - t / cx for offset 0.  Corners are numbered in the
// brick's node order: (0,0,0) (1,0,0) (1,1,0) (0,1,0), then z = 1.
// --------------------------------------------------------------------------
struct CCGrid {
  int W, H, D, cx, cy, cz, Wc, Hc, Dc;
  long ncell, nc, ncp;
};

__device__ __forceinline__ int cc_corner(int dx, int dy, int dz)
{
  return 4 * dz + (dy ? (dx ? 2 : 3) : (dx ? 1 : 0));
}
__device__ __forceinline__ void cc_offsets(int m, int &dx, int &dy, int &dz)
{
  dz = m >> 2;
  const int q = m & 3;
  dx = (q == 1 || q == 2);
  dy = (q >= 2);
}
__device__ __forceinline__ int cc_cell(int i, int f, int Nc)
{
  const int c = i / f;
  return c > Nc - 2 ? Nc - 2 : c;
}

// K_cell = sum over the cell's elements p of lambda_p G_p^l + mu_p G_p^m
__global__ __launch_bounds__(WG) void k_cc_galerkin(CCGrid g, const double *__restrict__ G,
                                                    const double *__restrict__ E, const double *__restrict__ nu,
                                                    double *__restrict__ KC)
{
  __shared__ double lam[UGK_CC_MAX_CELL], mu[UGK_CC_MAX_CELL];
  const long cell = blockIdx.x;
  const int ci = (int)(cell % (g.Wc - 1)), cj = (int)((cell / (g.Wc - 1)) % (g.Hc - 1));
  const int ck = (int)(cell / ((long)(g.Wc - 1) * (g.Hc - 1)));
  const int np = g.cx * g.cy * g.cz;
  const long nx = g.W - 1, ny = g.H - 1;
  for (int p = threadIdx.x; p < np; p += WG) {
    const int pi = p % g.cx, pj = (p / g.cx) % g.cy, pk = p / (g.cx * g.cy);
    const long e = (long)(ci * g.cx + pi) + nx * ((long)(cj * g.cy + pj) + ny * (long)(ck * g.cz + pk));
    const double Ee = E[e], v = nu[e];
    lam[p] = Ee * v / ((1.0 + v) * (1.0 - 2.0 * v));
    mu[p] = Ee / (2.0 * (1.0 + v));
  }
  __syncthreads();
  for (int o = threadIdx.x; o < 576; o += WG) {
    double a = 0.0;
    for (int p = 0; p < np; p++)
      a += lam[p] * G[(size_t)(2 * p) * 576 + o] + mu[p] * G[(size_t)(2 * p + 1) * 576 + o];
    KC[cell * 576 + o] = a;
  }
}

// A = 0 with the padding's diagonal 1
__global__ __launch_bounds__(WG) void k_cc_clear(double *A, long nc, long ncp)
{
  const long tot = ncp * ncp;
  for (long q = (long)blockIdx.x * WG + threadIdx.x; q < tot; q += (long)gridDim.x * WG) {
    const long i = q / ncp, j = q % ncp;
    A[q] = (i == j && i >= nc) ? 1.0 : 0.0;
  }
}

// A's blocks between coarse node I and its 27 neighbours: the sum over the
// cells holding both, in a fixed order; identity rows/columns where masked
__global__ __launch_bounds__(WG) void k_cc_assemble(CCGrid g, const double *__restrict__ KC,
                                                    const double *__restrict__ cm, double *A)
{
  const long nn = (long)g.Wc * g.Hc * g.Dc;
  const long t = (long)blockIdx.x * WG + threadIdx.x;
  if (t >= nn * 243) return;
  const long I = t / 243;
  const int rem = (int)(t % 243), nb = rem / 9, a = (rem % 9) / 3, b = rem % 3;
  const int Ii = (int)(I % g.Wc), Ij = (int)((I / g.Wc) % g.Hc), Ik = (int)(I / ((long)g.Wc * g.Hc));
  const int Ji = Ii + nb % 3 - 1, Jj = Ij + (nb / 3) % 3 - 1, Jk = Ik + nb / 9 - 1;
  if (Ji < 0 || Ji >= g.Wc || Jj < 0 || Jj >= g.Hc || Jk < 0 || Jk >= g.Dc) return;
  const long J = Ji + (long)g.Wc * (Jj + (long)g.Hc * Jk);
  const long row = 3 * I + a, col = 3 * J + b;
  double v = 0.0;
  if (cm[row] != 0.0 && cm[col] != 0.0) {
    const int k0 = max(max(Ik, Jk) - 1, 0), k1 = min(min(Ik, Jk), g.Dc - 2);
    const int j0 = max(max(Ij, Jj) - 1, 0), j1 = min(min(Ij, Jj), g.Hc - 2);
    const int i0 = max(max(Ii, Ji) - 1, 0), i1 = min(min(Ii, Ji), g.Wc - 2);
    for (int ck = k0; ck <= k1; ck++)
      for (int cj = j0; cj <= j1; cj++)
        for (int ci = i0; ci <= i1; ci++) {
          const int mI = cc_corner(Ii - ci, Ij - cj, Ik - ck), mJ = cc_corner(Ji - ci, Jj - cj, Jk - ck);
          const long cell = ci + (long)(g.Wc - 1) * (cj + (long)(g.Hc - 1) * ck);
          v += KC[cell * 576 + (3 * mI + a) * 24 + 3 * mJ + b];
        }
  } else if (row == col) {
    v = 1.0;
  }
  A[row * g.ncp + col] = v;
}

// ---- block Gauss-Jordan inversion of an SPD matrix, blocks of B -----------
// the pivot block's inverse, scalar Gauss-Jordan in LDS
template <int B>
__global__ __launch_bounds__(1024) void k_gj_pivot(const double *__restrict__ A, long ld, long K0,
                                                   double *__restrict__ piv)
{
  constexpr int PE = (B * B + 1023) / 1024;
  __shared__ double a[B][B + 1];
  for (int q = threadIdx.x; q < B * B; q += 1024) a[q / B][q % B] = A[(K0 + q / B) * ld + K0 + q % B];
  __syncthreads();
  for (int k = 0; k < B; k++) {
    const double p = 1.0 / a[k][k];
    double v[PE];
#pragma unroll
    for (int e = 0; e < PE; e++) {
      const int q = threadIdx.x + 1024 * e, i = q / B, j = q % B;
      if (q >= B * B) v[e] = 0.0;
      else if (i == k && j == k) v[e] = p;
      else if (i == k) v[e] = a[k][j] * p;
      else if (j == k) v[e] = -a[i][k] * p;
      else v[e] = a[i][j] - a[i][k] * a[k][j] * p;
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < PE; e++) {
      const int q = threadIdx.x + 1024 * e;
      if (q < B * B) a[q / B][q % B] = v[e];
    }
    __syncthreads();
  }
  for (int q = threadIdx.x; q < B * B; q += 1024) piv[q] = a[q / B][q % B];
}

// for the 64 columns J0.. of this block: R[:, J0..] = piv * A[K, J0..]
// (B x n), and for the same 64 rows Cb[J0.., :] = A[J0.., K] (n x B),
// copies taken before the in-place update; (B / 16) x 4 outputs per thread
template <int B>
__global__ __launch_bounds__(WG) void k_gj_panel(const double *__restrict__ A, long n, long K0,
                                                 const double *__restrict__ piv, double *__restrict__ R,
                                                 double *__restrict__ Cb)
{
  constexpr int RA = B / 16;
  __shared__ double P[B][B + 1];
  __shared__ double Ak[B][64 + 1];
  const long J0 = (long)blockIdx.x * 64;
  for (int q = threadIdx.x; q < B * B; q += WG) P[q / B][q % B] = piv[q];
  for (int q = threadIdx.x; q < B * 64; q += WG) Ak[q / 64][q % 64] = A[(K0 + q / 64) * n + J0 + q % 64];
  for (int q = threadIdx.x; q < 64 * B; q += WG) {
    const long j = J0 + q / B;
    Cb[j * B + q % B] = A[j * n + K0 + q % B];
  }
  __syncthreads();
  const int tc = threadIdx.x % 16, tr = threadIdx.x / 16;
  double acc[RA][4];
#pragma unroll
  for (int a = 0; a < RA; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) acc[a][b] = 0.0;
#pragma unroll 4
  for (int s = 0; s < B; s++) {
    double pa[RA], kb[4];
#pragma unroll
    for (int a = 0; a < RA; a++) pa[a] = P[tr + 16 * a][s];
#pragma unroll
    for (int b = 0; b < 4; b++) kb[b] = Ak[s][tc + 16 * b];
#pragma unroll
    for (int a = 0; a < RA; a++)
#pragma unroll
      for (int b = 0; b < 4; b++) acc[a][b] += pa[a] * kb[b];
  }
#pragma unroll
  for (int a = 0; a < RA; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) R[(long)(tr + 16 * a) * n + J0 + tc + 16 * b] = acc[a][b];
}

// one 64 x 64 tile of the update: A_ij -= C_iK R_Kj off the pivot rows and
// columns, A_iK = -C_iK piv, A_Kj = R_Kj, A_KK = piv.  The tile's R block is
// staged with piv in place of R at the pivot columns, so that one outer
// product serves both off-pivot cases and the pivot rows copy it; each
// thread owns 4 x 4 outputs (rows tr + 16 a, columns tc + 16 b)
template <int B>
__global__ __launch_bounds__(WG) void k_gj_update(double *__restrict__ A, long n, long K0,
                                                  const double *__restrict__ piv, const double *__restrict__ R,
                                                  const double *__restrict__ Cb)
{
  __shared__ double Cs[64][B + 1];
  __shared__ double Rs[B][64 + 1];
  const long I0 = (long)blockIdx.y * 64, J0 = (long)blockIdx.x * 64;
  for (int q = threadIdx.x; q < 64 * B; q += WG) Cs[q / B][q % B] = Cb[(I0 + q / B) * B + q % B];
  for (int q = threadIdx.x; q < B * 64; q += WG) {
    const int t = q / 64, c = q % 64;
    const long j = J0 + c;
    Rs[t][c] = (j >= K0 && j < K0 + B) ? piv[t * B + (j - K0)] : R[(long)t * n + j];
  }
  __syncthreads();
  const int tc = threadIdx.x % 16, tr = threadIdx.x / 16;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) acc[a][b] = 0.0;
#pragma unroll 4
  for (int t = 0; t < B; t++) {
    double ca[4], rb[4];
#pragma unroll
    for (int a = 0; a < 4; a++) ca[a] = Cs[tr + 16 * a][t];
#pragma unroll
    for (int b = 0; b < 4; b++) rb[b] = Rs[t][tc + 16 * b];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++) acc[a][b] += ca[a] * rb[b];
  }
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const int li = tr + 16 * a;
    const long i = I0 + li;
    const bool inI = i >= K0 && i < K0 + B;
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const int lj = tc + 16 * b;
      const long j = J0 + lj;
      const bool inJ = j >= K0 && j < K0 + B;
      double v;
      if (inI) v = Rs[i - K0][lj];
      else if (inJ) v = -acc[a][b];
      else v = A[i * n + j] - acc[a][b];
      A[i * n + j] = v;
    }
  }
}

// restriction, part 1: per cell, the weighted sums of its own fine nodes'
// masked residuals for each of its 8 corners x 3 directions
__global__ __launch_bounds__(WG) void k_cc_restrict(CCGrid g, const double *__restrict__ r,
                                                    const double *__restrict__ fm, double *cpart,
                                                    const int *stop)
{
  if (*stop) return;
  const long cell = blockIdx.x;
  const int ci = (int)(cell % (g.Wc - 1)), cj = (int)((cell / (g.Wc - 1)) % (g.Hc - 1));
  const int ck = (int)(cell / ((long)(g.Wc - 1) * (g.Hc - 1)));
  const int ni = g.cx + (ci == g.Wc - 2), nj = g.cy + (cj == g.Hc - 2), nk = g.cz + (ck == g.Dc - 2);
  const int tot = ni * nj * nk;
  double acc[24];
#pragma unroll
  for (int q = 0; q < 24; q++) acc[q] = 0.0;
  for (int q = threadIdx.x; q < tot; q += WG) {
    const int li = q % ni, lj = (q / ni) % nj, lk = q / (ni * nj);
    const long node = (long)(ci * g.cx + li) + (long)g.W * ((long)(cj * g.cy + lj) + (long)g.H * (ck * g.cz + lk));
    const double wx1 = (double)li / g.cx, wy1 = (double)lj / g.cy, wz1 = (double)lk / g.cz;
    double rv[3];
#pragma unroll
    for (int a = 0; a < 3; a++) rv[a] = fm[3 * node + a] * r[3 * node + a];
#pragma unroll
    for (int m = 0; m < 8; m++) {
      int dx, dy, dz;
      cc_offsets(m, dx, dy, dz);
      const double w = (dx ? wx1 : 1.0 - wx1) * (dy ? wy1 : 1.0 - wy1) * (dz ? wz1 : 1.0 - wz1);
#pragma unroll
      for (int a = 0; a < 3; a++) acc[3 * m + a] += w * rv[a];
    }
  }
  block_partials<24>(acc, cpart);
}

// restriction, part 2: g = cm * (sum of the adjacent cells' corner parts)
__global__ __launch_bounds__(WG) void k_cc_gather(CCGrid g, const double *__restrict__ cpart,
                                                  const double *__restrict__ cm, double *__restrict__ gv,
                                                  const int *stop)
{
  if (*stop) return;
  const long t = (long)blockIdx.x * WG + threadIdx.x;
  if (t >= g.ncp) return;
  if (t >= g.nc) { gv[t] = 0.0; return; }
  const long I = t / 3;
  const int a = (int)(t % 3);
  const int Ii = (int)(I % g.Wc), Ij = (int)((I / g.Wc) % g.Hc), Ik = (int)(I / ((long)g.Wc * g.Hc));
  double s = 0.0;
  for (int ck = max(Ik - 1, 0); ck <= min(Ik, g.Dc - 2); ck++)
    for (int cj = max(Ij - 1, 0); cj <= min(Ij, g.Hc - 2); cj++)
      for (int ci = max(Ii - 1, 0); ci <= min(Ii, g.Wc - 2); ci++) {
        const long cell = ci + (long)(g.Wc - 1) * (cj + (long)(g.Hc - 1) * ck);
        const int m = cc_corner(Ii - ci, Ij - cj, Ik - ck);
        s += cpart[(long)(3 * m + a) * g.ncell + cell];
      }
  gv[t] = cm[t] * s;
}

// y = A^-1 g, one wavefront per row
__global__ __launch_bounds__(WG) void k_cc_gemv(const double *__restrict__ A, long n, const double *__restrict__ gv,
                                                double *__restrict__ y, const int *stop)
{
  if (stop && *stop) return;
  const int lane = threadIdx.x & 63;
  const long row = (long)blockIdx.x * NW + (threadIdx.x >> 6);
  if (row >= n) return;
  // n is even (padded to a multiple of 64), so rows and g split into
  // 16-byte pairs; two pairs per lane in flight, two accumulators
  typedef Pair<double>::type D2;
  const D2 *a = (const D2 *)(A + row * n), *g2 = (const D2 *)gv;
  const long np = n >> 1;
  double s0 = 0.0, s1 = 0.0;
  long p = lane;
  for (; p + 64 < np; p += 128) {
    const D2 x0 = a[p], x1 = a[p + 64], h0 = g2[p], h1 = g2[p + 64];
    s0 += x0.x * h0.x + x0.y * h0.y;
    s1 += x1.x * h1.x + x1.y * h1.y;
  }
  if (p < np) {
    const D2 x0 = a[p], h0 = g2[p];
    s0 += x0.x * h0.x + x0.y * h0.y;
  }
  const double s = wave_sum(s0 + s1);
  if (lane == 0) y[row] = s;
}

// prolongation: z_c = fm P y; MODE 1: z = dinv r + z_c with (r, z) partials
// for the PCG direction update, MODE 0: z_c alone
template <int MODE>
__global__ __launch_bounds__(WG) void k_cc_z(CCGrid g, const double *__restrict__ y, const double *__restrict__ fm,
                                             const double *__restrict__ r, const double *__restrict__ dinv,
                                             double *__restrict__ z, const int *stop, double *part)
{
  if (*stop) return;
  const long nn = (long)g.W * g.H * g.D;
  double v[1] = {0.0};
  long n1, n2;
  xcd_chunk(nn, n1, n2);
  for (long node = n1 + threadIdx.x; node < n2; node += WG) {
    const int i = (int)(node % g.W), j = (int)((node / g.W) % g.H), k = (int)(node / ((long)g.W * g.H));
    const int ci = cc_cell(i, g.cx, g.Wc), cj = cc_cell(j, g.cy, g.Hc), ck = cc_cell(k, g.cz, g.Dc);
    const double wx1 = (double)(i - ci * g.cx) / g.cx, wy1 = (double)(j - cj * g.cy) / g.cy;
    const double wz1 = (double)(k - ck * g.cz) / g.cz;
    double zc[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int m = 0; m < 8; m++) {
      int dx, dy, dz;
      cc_offsets(m, dx, dy, dz);
      const double w = (dx ? wx1 : 1.0 - wx1) * (dy ? wy1 : 1.0 - wy1) * (dz ? wz1 : 1.0 - wz1);
      const long I = (ci + dx) + (long)g.Wc * ((cj + dy) + (long)g.Hc * (ck + dz));
#pragma unroll
      for (int a = 0; a < 3; a++) zc[a] += w * y[3 * I + a];
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const long d = 3 * node + a;
      const double c = fm[d] * zc[a];
      if (MODE) {
        const double zi = dinv[d] * r[d] + c;
        z[d] = zi;
        v[0] += r[d] * zi;
      } else {
        z[d] = c;
      }
    }
  }
  if (MODE) block_partials<1>(v, PART_S(part));
}

static inline hipStream_t S(void *s) { return (hipStream_t)s; }
// blocks for one slice per wavefront, rounded up to a multiple of the 8 XCDs
// so that remap_block() is a bijection on [0, gridDim.x)
static inline unsigned spmv_blocks(const ugk_mat *K)
{
  long b = (K->n_slices + 1) / 2;       // two slices per 4-wave block (slot-split)
  return (unsigned)((b + 7) / 8 * 8);
}
// blocks for kernels that still map one slice per wavefront (k_pcg_a, FP32 inner)
static inline unsigned spmv_blocks1(const ugk_mat *K)
{
  long b = (K->n_slices + NW - 1) / NW;
  return (unsigned)((b + 7) / 8 * 8);
}
static int g_ncu = 256;
static int g_vec_cap = 512;   // blocks of the vector kernels (UFEM_VEC_BLOCKS)
static int g_gj_block = 32;   // block size of the dense Gauss-Jordan inverse (UFEM_GJ_BLOCK: 32 or 64)
// Vector kernels also sum the previous kernel's block partials redundantly in
// every block, so their grid is kept small: 2 blocks per CU stream at ~95% of
// the bandwidth of larger grids while reading 4x fewer partials.
static inline unsigned vec_blocks(long n)
{
  long b = (n + WG - 1) / WG;
  long cap = g_vec_cap;
  return (unsigned)(b < 1 ? 1 : (b > cap ? cap : b));
}
static const unsigned RED_MAX_BLOCKS = UGK_RED_MAX_BLOCKS;

// Plain streaming read of the matrix's values and columns: 16-byte loads,
// grid-stride, one load in flight per lane per step (NT: non-temporal, as the
// SpMV issues them), each wave folding what it read into one word of `out`.
// It moves the SpMV's matrix bytes with none of its gathers or row sums: the
// read rate that stream can reach on this device.
template <bool NT>
__global__ void __launch_bounds__(WG) k_stream_read(const double *__restrict__ a, long na,
                                                     const double *__restrict__ b, long nb, double *out)
{
  typedef Pair<double>::type D2;
  const long stride = (long)gridDim.x * WG;
  double s = 0;
  for (long i = blockIdx.x * (long)WG + threadIdx.x; i < na; i += stride) {
    const D2 v = NT ? ldnt((const D2 *)a + i) : ((const D2 *)a)[i];
    s += v.x + v.y;
  }
  for (long i = blockIdx.x * (long)WG + threadIdx.x; i < nb; i += stride) {
    const D2 v = NT ? ldnt((const D2 *)b + i) : ((const D2 *)b)[i];
    s += v.x + v.y;
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) out[blockIdx.x * NW + (threadIdx.x >> 6)] = s;
}

extern "C" {

int ugk_init(void)
{
  int dev = 0, n = 0;
  CK(hipGetDeviceCount(&n));
  if (n < 1) { snprintf(g_err, sizeof(g_err), "no HIP device"); return -1; }
  CK(hipGetDevice(&dev));
  hipDeviceProp_t pr;
  CK(hipGetDeviceProperties(&pr, dev));
  g_ncu = pr.multiProcessorCount;
  g_vec_cap = 2 * g_ncu;
  if (const char *e = getenv("UFEM_VEC_BLOCKS")) {
    int v = atoi(e);
    if (v >= 8 && v <= 16384) g_vec_cap = v;
  }
  if (const char *e = getenv("UFEM_GJ_BLOCK")) g_gj_block = atoi(e) == 64 ? 64 : 32;
  return 0;
}

const char *ugk_error(void) { return g_err; }

void *ugk_malloc(size_t bytes)
{
  void *p = nullptr;
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) { fail("hipMalloc", e); return nullptr; }
  // hipMemset runs on the null stream, which does not order itself with the
  // library's non-blocking streams: finish it before anyone uploads into p
  e = hipMemset(p, 0, bytes);
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) { fail("hipMemset", e); (void)hipFree(p); return nullptr; }
  return p;
}
void ugk_free(void *p)
{
  if (p) (void)hipFree(p);
}
int ugk_h2d(void *d, const void *h, size_t bytes, void *s)
{
  CK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, S(s)));
  return 0;
}
int ugk_d2h(void *h, const void *d, size_t bytes, void *s)
{
  CK(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, S(s)));
  CK(hipStreamSynchronize(S(s)));
  return 0;
}
int ugk_d2d(void *d, const void *src, size_t bytes, void *s)
{
  CK(hipMemcpyAsync(d, src, bytes, hipMemcpyDeviceToDevice, S(s)));
  return 0;
}
int ugk_zero(void *d, size_t bytes, void *s)
{
  long w = (long)(bytes / 4);
  if (w == 0) return 0;
  k_fill_zero<<<vec_blocks(w), WG, 0, S(s)>>>(w, (unsigned int *)d);
  CK_LAUNCH("k_fill_zero");
  return 0;
}
void *ugk_stream_create(void)
{
  hipStream_t s = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
  return (void *)s;
}
void ugk_stream_destroy(void *s)
{
  if (s) (void)hipStreamDestroy(S(s));
}
int ugk_sync(void *s)
{
  CK(hipStreamSynchronize(S(s)));
  return 0;
}
void *ugk_event_create(void)
{
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return (void *)e;
}
void ugk_event_destroy(void *e)
{
  if (e) (void)hipEventDestroy((hipEvent_t)e);
}
int ugk_event_record(void *e, void *s)
{
  CK(hipEventRecord((hipEvent_t)e, S(s)));
  return 0;
}
float ugk_event_ms(void *e0, void *e1)
{
  float ms = -1.0f;
  if (hipEventSynchronize((hipEvent_t)e1) != hipSuccess) return -1.0f;
  if (hipEventElapsedTime(&ms, (hipEvent_t)e0, (hipEvent_t)e1) != hipSuccess) return -1.0f;
  return ms;
}
int ugk_device_info(int *n_cu, long *hbm, char *name, int name_len)
{
  int dev = 0;
  hipDeviceProp_t pr;
  CK(hipGetDevice(&dev));
  CK(hipGetDeviceProperties(&pr, dev));
  if (n_cu) *n_cu = pr.multiProcessorCount;
  if (hbm) *hbm = (long)pr.totalGlobalMem;
  if (name && name_len > 0) snprintf(name, (size_t)name_len, "%s (%s)", pr.name, pr.gcnArchName);
  return 0;
}

void *ugk_host_alloc(size_t bytes)
{
  void *p = nullptr;
  hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
  if (e != hipSuccess) { fail("hipHostMalloc", e); return nullptr; }
  memset(p, 0, bytes);
  return p;
}
void ugk_host_free(void *p)
{
  if (p) (void)hipHostFree(p);
}
void *ugk_host_alloc_mapped(size_t bytes, void **dptr)
{
  void *p = nullptr;
  hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocMapped | hipHostMallocCoherent);
  if (e != hipSuccess) { fail("hipHostMalloc(mapped)", e); return nullptr; }
  memset(p, 0, bytes);
  e = hipHostGetDevicePointer(dptr, p, 0);
  if (e != hipSuccess) { fail("hipHostGetDevicePointer", e); (void)hipHostFree(p); return nullptr; }
  return p;
}
int ugk_event_done(void *e)
{
  hipError_t r = hipEventQuery((hipEvent_t)e);
  if (r == hipSuccess) return 1;
  if (r == hipErrorNotReady) return 0;
  fail("hipEventQuery", r);
  return -1;
}
int ugk_d2h_async(void *h, const void *d, size_t bytes, void *s)
{
  CK(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, S(s)));
  return 0;
}
int ugk_event_sync(void *e)
{
  CK(hipEventSynchronize((hipEvent_t)e));
  return 0;
}

int ugk_graph_begin(void *s)
{
  CK(hipStreamBeginCapture(S(s), hipStreamCaptureModeThreadLocal));
  return 0;
}
void *ugk_graph_end(void *s)
{
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  hipError_t e = hipStreamEndCapture(S(s), &g);
  if (e != hipSuccess) { fail("hipStreamEndCapture", e); return nullptr; }
  e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) { fail("hipGraphInstantiate", e); return nullptr; }
  return (void *)ge;
}
int ugk_graph_launch(void *g, void *s)
{
  CK(hipGraphLaunch((hipGraphExec_t)g, S(s)));
  return 0;
}
void ugk_graph_destroy(void *g)
{
  if (g) (void)hipGraphExecDestroy((hipGraphExec_t)g);
}

// ---- assembly -------------------------------------------------------------
int ugk_elem_results(long ne, int npts, const double *d_x, const double *d_y, const double *d_z,
                     const int *d_enodes, const double *d_E, const double *d_nu, const double *d_u, double *d_out,
                     ugk_state *d_st, void *s)
{
  const unsigned nb = (unsigned)((ne + NW - 1) / NW);
  if (ne <= 0) return 0;
  if (npts == 2) k_elem_results<2><<<nb, WG, 0, S(s)>>>(ne, d_x, d_y, d_z, d_enodes, d_E, d_nu, d_u, d_out, d_st);
  else if (npts == 3) k_elem_results<3><<<nb, WG, 0, S(s)>>>(ne, d_x, d_y, d_z, d_enodes, d_E, d_nu, d_u, d_out, d_st);
  else { snprintf(g_err, sizeof(g_err), "npts must be 2 or 3"); return -1; }
  CK_LAUNCH("k_elem_results");
  return 0;
}

int ugk_elem_stiff_dense(long n_elems, const double *d_xyz, const int *d_enodes, const double *d_E,
                         const double *d_nu, int npts, double *d_Ke, ugk_state *d_st, void *s)
{
  (void)d_enodes;
  if (n_elems <= 0) return 0;
  SellView M{};
  unsigned nb = (unsigned)((n_elems + NW - 1) / NW);
  if (npts == 2)
    k_elem_stiff<2, true><<<nb, WG, 0, S(s)>>>(n_elems, nullptr, d_xyz, nullptr, nullptr, nullptr, d_E,
                                               d_nu, d_Ke, M, nullptr, d_st);
  else if (npts == 3)
    k_elem_stiff<3, true><<<nb, WG, 0, S(s)>>>(n_elems, nullptr, d_xyz, nullptr, nullptr, nullptr, d_E,
                                               d_nu, d_Ke, M, nullptr, d_st);
  else { snprintf(g_err, sizeof(g_err), "npts must be 2 or 3"); return -1; }
  CK_LAUNCH("k_elem_stiff dense");
  return 0;
}

int ugk_elem_stiff_scatter(const ugk_mat *K, long n_list, const int *d_elist, const double *d_x,
                           const double *d_y, const double *d_z, const int *d_enodes,
                           const double *d_E, const double *d_nu, int npts, ugk_state *d_st, void *s)
{
  if (n_list <= 0) return 0;
  if (K->bs != 3) { snprintf(g_err, sizeof(g_err), "scatter needs 3x3 blocks"); return -1; }
  unsigned nb = (unsigned)((n_list + NW - 1) / NW);
  if (npts == 2)
    k_elem_stiff<2, false><<<nb, WG, 0, S(s)>>>(n_list, d_elist, d_x, d_y, d_z, d_enodes, d_E, d_nu,
                                                nullptr, view(K), K->d_vals, d_st);
  else if (npts == 3)
    k_elem_stiff<3, false><<<nb, WG, 0, S(s)>>>(n_list, d_elist, d_x, d_y, d_z, d_enodes, d_E, d_nu,
                                                nullptr, view(K), K->d_vals, d_st);
  else { snprintf(g_err, sizeof(g_err), "npts must be 2 or 3"); return -1; }
  CK_LAUNCH("k_elem_stiff scatter");
  return 0;
}

int ugk_save_rows(const ugk_mat *K, long n_fixed, const int *d_fixed, int maxb, double *d_rows,
                  int *d_rcols, void *s)
{
  if (n_fixed <= 0) return 0;
  k_save_rows<<<(unsigned)((n_fixed + 127) / 128), 128, 0, S(s)>>>(view(K), K->d_vals, n_fixed, d_fixed,
                                                                    maxb, d_rows, d_rcols);
  CK_LAUNCH("k_save_rows");
  return 0;
}

int ugk_apply_zero_bc(const ugk_mat *K, long n_fixed, const int *d_fixed, void *s)
{
  if (n_fixed <= 0) return 0;
  unsigned nb = (unsigned)((n_fixed + 127) / 128);
  k_zero_rowcol<<<nb, 128, 0, S(s)>>>(view(K), K->d_vals, n_fixed, d_fixed);
  CK_LAUNCH("k_zero_rowcol");
  k_set_diag_one<<<nb, 128, 0, S(s)>>>(view(K), K->d_vals, n_fixed, d_fixed);
  CK_LAUNCH("k_set_diag_one");
  return 0;
}

int ugk_reactions(const ugk_mat *K, long n_fixed, const int *d_fixed, int maxb, const double *d_rows,
                  const int *d_rcols, const double *d_u, const double *d_Forig, double *d_R, void *s)
{
  (void)K;
  if (n_fixed <= 0) return 0;
  k_reactions<<<(unsigned)((n_fixed + 127) / 128), 128, 0, S(s)>>>(n_fixed, d_fixed, maxb, d_rows, d_rcols,
                                                                     d_u, d_Forig, d_R);
  CK_LAUNCH("k_reactions");
  return 0;
}

int ugk_vals_to_f32(const ugk_mat *K, void *s)
{
  long n = K->total_slots * 64L * K->bs * K->bs;
  if (!K->d_vals32) { snprintf(g_err, sizeof(g_err), "no fp32 buffer"); return -1; }
  if (K->bs == 3) k_vals_f32_b3<<<vec_blocks(n), WG, 0, S(s)>>>(n, K->d_vals, K->d_vals32);
  else k_vals_f32<float><<<vec_blocks(n), WG, 0, S(s)>>>(n, K->d_vals, K->d_vals32);
  CK_LAUNCH("k_vals_f32");
  if (K->d_rem16) {
    if (K->bs == 3) k_vals_rem16<true><<<vec_blocks(n), WG, 0, S(s)>>>(n, K->d_vals, K->d_rem16);
    else k_vals_rem16<false><<<vec_blocks(n), WG, 0, S(s)>>>(n, K->d_vals, K->d_rem16);
    CK_LAUNCH("k_vals_rem16");
  }
  if (K->bs == 3 && K->d_corr32) {
    k_corr32<<<(unsigned)((K->n_brows + WG - 1) / WG), WG, 0, S(s)>>>(view(K), K->d_vals, K->d_rem16, K->d_corr32);
    CK_LAUNCH("k_corr32");
  }
  return 0;
}

int ugk_scatter_vals(long n, const int *d_map, const double *d_src, double *d_dst, void *s)
{
  if (n <= 0) return 0;
  k_scatter_vals<<<vec_blocks(n), WG, 0, S(s)>>>(n, d_map, d_src, d_dst);
  CK_LAUNCH("k_scatter_vals");
  return 0;
}

// ---- sparse ---------------------------------------------------------------
int ugk_stream_read(const ugk_mat *K, int nontemporal, double *d_out, void *s)
{
  const long slots = K->total_slots * 64;                  // a multiple of 64
  const long na = slots * K->bs * K->bs / 2, nb = slots / 4;   // 16-byte words (columns: 4 per word)
  const unsigned blocks = (unsigned)(g_ncu * 8);
  if ((long)blocks * NW > UGK_RED_MAX_BLOCKS) return -1;
  if (nontemporal)
    k_stream_read<true><<<blocks, WG, 0, S(s)>>>(K->d_vals, na, (const double *)K->d_cols, nb, d_out);
  else
    k_stream_read<false><<<blocks, WG, 0, S(s)>>>(K->d_vals, na, (const double *)K->d_cols, nb, d_out);
  CK_LAUNCH("k_stream_read");
  return 0;
}
int ugk_spmv(const ugk_mat *K, const double *d_x, double *d_y, void *s)
{
  if (K->bs == 3) k_spmv<3, double, double><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, d_x, d_y);
  else k_spmv<1, double, double><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, d_x, d_y);
  CK_LAUNCH("k_spmv");
  return 0;
}
int ugk_spmv_f32x64(const ugk_mat *K, const double *d_x, double *d_y, void *s)
{
  if (!K->d_vals32) { snprintf(g_err, sizeof(g_err), "no fp32 copy of the matrix"); return -1; }
  if (K->bs == 3) k_spmv<3, float, double><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals32, d_x, d_y);
  else k_spmv<1, float, double><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals32, d_x, d_y);
  CK_LAUNCH("k_spmv_f32x64");
  return 0;
}
int ugk_spmv_f32(const ugk_mat *K, const float *d_x, float *d_y, void *s)
{
  if (K->bs == 3) k_spmv<3, float, float><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals32, d_x, d_y);
  else k_spmv<1, float, float><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals32, d_x, d_y);
  CK_LAUNCH("k_spmv_f32");
  return 0;
}
int ugk_diag_inv(const ugk_mat *K, double *d_dinv, ugk_state *d_st, void *s)
{
  long n = K->n_brows * K->bs;
  unsigned nb = (unsigned)((n + WG - 1) / WG);
  if (K->bs == 3) k_diag_inv<3><<<nb, WG, 0, S(s)>>>(view(K), K->d_vals, d_dinv, d_st);
  else k_diag_inv<1><<<nb, WG, 0, S(s)>>>(view(K), K->d_vals, d_dinv, d_st);
  CK_LAUNCH("k_diag_inv");
  return 0;
}
int ugk_mat_norm(const ugk_mat *K, ugk_state *d_st, double *d_part, void *s)
{
  long n = K->n_brows * K->bs;
  if (K->bs == 3) k_mat_norm<3><<<vec_blocks(n), WG, 0, S(s)>>>(view(K), K->d_vals, d_st, d_part);
  else k_mat_norm<1><<<vec_blocks(n), WG, 0, S(s)>>>(view(K), K->d_vals, d_st, d_part);
  CK_LAUNCH("k_mat_norm");
  return 0;
}

int ugk_dot(long n, const double *a, const double *b, double *out, double *part, void *s)
{
  // the counter lives after the partials (part holds UGK_PART_DOUBLES)
  unsigned int *ctr = (unsigned int *)(part + 8L * RED_MAX_BLOCKS);
  k_dot<<<vec_blocks(n), WG, 0, S(s)>>>(n, a, b, out, part, ctr);
  CK_LAUNCH("k_dot");
  return 0;
}
int ugk_axpby(long n, double a, const double *x, double b, const double *y, double *z, void *s)
{
  k_axpby<<<vec_blocks(n), WG, 0, S(s)>>>(n, a, x, b, y, z);
  CK_LAUNCH("k_axpby");
  return 0;
}

// ---- PCG ------------------------------------------------------------------
int ugk_pcg_start(const ugk_mat *K, long n, const double *b, const double *x, double *r, double *z,
                  double *p0, double *p1, const double *dinv, ugk_state *st, double *part, void *s)
{
  (void)n;
  (void)st;
  if (K->bs == 3)
    k_pcg_start<3><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, b, x, r, z, p0, p1, dinv, part);
  else
    k_pcg_start<1><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, b, x, r, z, p0, p1, dinv, part);
  CK_LAUNCH("k_pcg_start");
  return 0;
}

int ugk_pcg_iter(const ugk_mat *K, long n, double *x, double *r, double *z, double **p, double *q,
                 const double *dinv, ugk_state *st, double *part, int variant, void *s)
{
  const unsigned GS = spmv_blocks(K), GB = vec_blocks(n);
  const unsigned GA = variant == 1 ? spmv_blocks(K) : spmv_blocks1(K);
  if (variant == 1) {
    k_pcg_p<<<GB, WG, 0, S(s)>>>(n, z, p[0], st, part, GS, GB);
    CK_LAUNCH("k_pcg_p");
    if (K->bs == 3) k_pcg_q<3><<<GA, WG, 0, S(s)>>>(view(K), K->d_vals, p[0], q, st, part);
    else k_pcg_q<1><<<GA, WG, 0, S(s)>>>(view(K), K->d_vals, p[0], q, st, part);
    CK_LAUNCH("k_pcg_q");
    k_pcg_b<<<GB, WG, 0, S(s)>>>(n, x, r, z, p[0], p[0], q, dinv, st, part, GA, 0);
    CK_LAUNCH("k_pcg_b");
  } else {
    if (K->bs == 3) k_pcg_a<3><<<GA, WG, 0, S(s)>>>(view(K), K->d_vals, z, p[0], p[1], q, st, part, GS, GB);
    else k_pcg_a<1><<<GA, WG, 0, S(s)>>>(view(K), K->d_vals, z, p[0], p[1], q, st, part, GS, GB);
    CK_LAUNCH("k_pcg_a");
    k_pcg_b<<<GB, WG, 0, S(s)>>>(n, x, r, z, p[0], p[1], q, dinv, st, part, GA, 1);
    CK_LAUNCH("k_pcg_b");
  }
  return 0;
}

int ugk_pcg_iter_a(const ugk_mat *K, const double *z, const double *pold, double *pnew, double *q,
                   ugk_state *st, double *part, void *s)
{
  // the fused kernel as a solve runs it, from a state with iter = 1
  (void)pold;
  const unsigned GA = spmv_blocks1(K);
  if (K->bs == 3)
    k_pcg_a<3><<<GA, WG, 0, S(s)>>>(view(K), K->d_vals, z, pnew, (double *)pold, q, st, part, GA, GA);
  else
    k_pcg_a<1><<<GA, WG, 0, S(s)>>>(view(K), K->d_vals, z, pnew, (double *)pold, q, st, part, GA, GA);
  CK_LAUNCH("k_pcg_a");
  return 0;
}

int ugk_cg1_start(const ugk_mat *K, long n, const double *b, const double *x, double *r, double *u, double *p,
                  double *sv, const double *dinv, double *part, void *s)
{
  (void)n;
  // the start's partials share the parity slot sized for vector grids
  if (spmv_blocks(K) > 16384) { snprintf(g_err, sizeof(g_err), "mesh too large for the single-reduction form"); return -1; }
  if (K->bs == 3)
    k_cg1_start<3><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, b, x, r, u, p, sv, dinv, part);
  else
    k_cg1_start<1><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, b, x, r, u, p, sv, dinv, part);
  CK_LAUNCH("k_cg1_start");
  return 0;
}

int ugk_cg1_iter(const ugk_mat *K, long n, double *x, double *r, double *u, double *p, double *sv, double *w,
                 const double *dinv, ugk_state *st, double *part, void *s)
{
  const unsigned G0 = spmv_blocks(K), GA = spmv_blocks(K), GB = vec_blocks(n);
  if (K->bs == 3) k_cg1_w<3><<<GA, WG, 0, S(s)>>>(view(K), K->d_vals, u, w, st, part);
  else k_cg1_w<1><<<GA, WG, 0, S(s)>>>(view(K), K->d_vals, u, w, st, part);
  CK_LAUNCH("k_cg1_w");
  k_cg1_upd<<<GB, WG, 0, S(s)>>>(n, x, r, u, p, sv, w, dinv, st, part, G0, GB, GA);
  CK_LAUNCH("k_cg1_upd");
  return 0;
}

int ugk_cg1_iter_w(const ugk_mat *K, const double *u, double *w, ugk_state *st, double *part, void *s)
{
  if (K->bs == 3) k_cg1_w<3><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, u, w, st, part);
  else k_cg1_w<1><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, u, w, st, part);
  CK_LAUNCH("k_cg1_w");
  return 0;
}

int ugk_pcg_iter_q(const ugk_mat *K, const double *p, double *q, ugk_state *st, double *part, void *s)
{
  if (K->bs == 3) k_pcg_q<3><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, p, q, st, part);
  else k_pcg_q<1><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, p, q, st, part);
  CK_LAUNCH("k_pcg_q");
  return 0;
}

int ugk_pcg_persistent(const ugk_mat *K, long n, double *x, double *r, double *z, double *p0,
                       double *p1, double *q, const double *dinv, ugk_state *st, double *part,
                       void *s)
{
  int per_cu = 0;
  const char *env = getenv("UFEM_GRID_SYNC");
  const bool cgs = env && env[0] == 'c';
  const void *fn = K->bs == 3 ? (cgs ? (const void *)k_pcg_persistent<3, true> : (const void *)k_pcg_persistent<3, false>)
                              : (cgs ? (const void *)k_pcg_persistent<1, true> : (const void *)k_pcg_persistent<1, false>);
  CK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, WG, 0));
  if (per_cu < 1) { snprintf(g_err, sizeof(g_err), "persistent kernel does not fit"); return -1; }
  if (per_cu > 4) per_cu = 4;
  unsigned G = (unsigned)(per_cu * g_ncu);
  SellView M = view(K);
  const double *vals = K->d_vals;
  unsigned GS = spmv_blocks(K);
  if (G > 16384) G = 16384;
  /* the grid barrier gives up after about 10 s of the device wall clock */
  long limit = ugk_wall_clock_khz() * 10000L;
  if (limit <= 0) limit = 1L << 40;
  void *args[] = {(void *)&M, (void *)&vals, (void *)&n, (void *)&x, (void *)&r, (void *)&z,
                  (void *)&p0, (void *)&p1, (void *)&q, (void *)&dinv, (void *)&st, (void *)&part,
                  (void *)&GS, (void *)&limit};
  CK(hipLaunchCooperativeKernel(fn, dim3(G), dim3(WG), args, 0, S(s)));
  return 0;
}

// ---- mixed precision -------------------------------------------------------
int ugk_mx_start(const ugk_mat *K, long n, const double *b, const double *x64, double *r64, float *r32,
                 float *z32, float *p32, float *x32, const float *dinv32, double *part, void *s)
{
  (void)n;
  if (K->bs == 3)
    k_mx_start<3><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, b, x64, r64, r32, z32, p32, x32, dinv32,
                                                   part);
  else
    k_mx_start<1><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, b, x64, r64, r32, z32, p32, x32, dinv32,
                                                   part);
  CK_LAUNCH("k_mx_start");
  return 0;
}

int ugk_mx_iter(const ugk_mat *K, long n, const double *b, double *x64, double *r64, float *r32, float *z32,
                float *p32, float *q32, float *x32, const float *dinv32, ugk_state *st, double *part,
                void *s)
{
  const unsigned GS = spmv_blocks(K), GA = spmv_blocks(K), GR = spmv_blocks(K), GB = vec_blocks(n);
  k_mx_p<<<GB, WG, 0, S(s)>>>(n, z32, p32, st, part, GS, GB, GR);
  CK_LAUNCH("k_mx_p");
  if (K->bs == 3) k_mx_q<3><<<GA, WG, 0, S(s)>>>(view(K), K->d_vals32, p32, q32, st, part);
  else k_mx_q<1><<<GA, WG, 0, S(s)>>>(view(K), K->d_vals32, p32, q32, st, part);
  CK_LAUNCH("k_mx_q");
  k_mx_b<<<GB, WG, 0, S(s)>>>(n, x32, r32, z32, p32, q32, dinv32, st, part, GA);
  CK_LAUNCH("k_mx_b");
  k_mx_ru<<<GB, WG, 0, S(s)>>>(n, x64, x32, st, part, GB);
  CK_LAUNCH("k_mx_ru");
  if (K->bs == 3) k_mx_r<3><<<GR, WG, 0, S(s)>>>(view(K), K->d_vals, b, x64, r64, r32, z32, dinv32, st, part);
  else k_mx_r<1><<<GR, WG, 0, S(s)>>>(view(K), K->d_vals, b, x64, r64, r32, z32, dinv32, st, part);
  CK_LAUNCH("k_mx_r");
  return 0;
}

int ugk_residual(const ugk_mat *K, long n, const double *b, const double *x, double *r, ugk_state *st,
                 double *part, void *s)
{
  (void)n;
  if (K->bs == 3)
    k_residual<3><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, b, x, r, st, part);
  else
    k_residual<1><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, b, x, r, st, part);
  CK_LAUNCH("k_residual");
  return 0;
}

int ugk_to_f32(long n, const double *a, float *b, void *s)
{
  k_vals_f32<float><<<vec_blocks(n), WG, 0, S(s)>>>(n, a, b);
  CK_LAUNCH("k_to_f32");
  return 0;
}

// ---- BiCGStab ---------------------------------------------------------------
int ugk_bicgs_start(const ugk_mat *K, long n, const double *b, const double *x, double *r, double *rhat,
                    double *p, double *v, ugk_state *st, double *part, void *s)
{
  if (ugk_spmv(K, x, r, s)) return -1;
  k_bicgs_init<<<vec_blocks(n), WG, 0, S(s)>>>(n, b, r, rhat, p, v, st, part);
  CK_LAUNCH("k_bicgs_init");
  return 0;
}
int ugk_bicgs_iter(const ugk_mat *K, long n, double *x, double *r, const double *rhat, double *p,
                   double *phat, double *v, double *sv, double *shat, double *t, const double *dinv,
                   ugk_state *st, double *part, void *s)
{
  k_bicgs_p<<<vec_blocks(n), WG, 0, S(s)>>>(n, r, p, v, phat, dinv, st);
  CK_LAUNCH("k_bicgs_p");
  if (K->bs == 3)
    k_spmv_dots<3, 1><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, phat, v, rhat, st, part, 1, 1);
  else
    k_spmv_dots<1, 1><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, phat, v, rhat, st, part, 1, 1);
  CK_LAUNCH("k_spmv_dots v");
  k_bicgs_s<<<vec_blocks(n), WG, 0, S(s)>>>(n, r, v, sv, shat, dinv, st, part);
  CK_LAUNCH("k_bicgs_s");
  if (K->bs == 3)
    k_spmv_dots<3, 2><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, shat, t, sv, st, part, 3, 3);
  else
    k_spmv_dots<1, 2><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, shat, t, sv, st, part, 3, 3);
  CK_LAUNCH("k_spmv_dots t");
  k_bicgs_x<<<vec_blocks(n), WG, 0, S(s)>>>(n, x, r, rhat, phat, sv, shat, t, st, part);
  CK_LAUNCH("k_bicgs_x");
  return 0;
}

// ---- batched solve -----------------------------------------------------------
int ugk_batch_occupancy(int wg)
{
  int per = 0;
  const void *fn = wg == 256 ? (const void *)k_batch_pcg<256>
                 : (wg == 512 ? (const void *)k_batch_pcg<512> : (const void *)k_batch_pcg<1024>);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, wg, 0) != hipSuccess) return 0;
  return per;
}

int ugk_batch_pcg(const ugk_bsample *samples, const int *slots, int nslots, unsigned *ticket, unsigned *tctr,
                  double *tpart, ugk_bresult *res, double tol, long maxit, int wg, int grid, const int *d_abort,
                  long wait_ticks, void *s)
{
  if (grid < 1) grid = g_ncu;
  const int2 *sl = (const int2 *)slots;
  BatchStop bs;
  bs.host = d_abort;
  bs.dev = ticket + 1;                 // zeroed with the ticket
  bs.limit = wait_ticks > 0 ? wait_ticks : (1L << 40);
  if (wg == 256) k_batch_pcg<256><<<grid, 256, 0, S(s)>>>(samples, sl, nslots, ticket, tctr, tpart, res, tol, maxit, bs);
  else if (wg == 512) k_batch_pcg<512><<<grid, 512, 0, S(s)>>>(samples, sl, nslots, ticket, tctr, tpart, res, tol, maxit, bs);
  else if (wg == 1024) k_batch_pcg<1024><<<grid, 1024, 0, S(s)>>>(samples, sl, nslots, ticket, tctr, tpart, res, tol, maxit, bs);
  else { snprintf(g_err, sizeof(g_err), "workgroup size %d not built", wg); return -1; }
  CK_LAUNCH("k_batch_pcg");
  return 0;
}

int ugk_diag_inv3(const ugk_mat *K, double *d_dinv, ugk_state *d_st, void *s)
{
  return ugk_diag_inv(K, d_dinv, d_st, s);
}

long ugk_wall_clock_khz(void)
{
  int dev = 0, khz = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess) return 0;
  return khz;
}

int ugk_ssor_permute_vals(const ugk_mat *K, const ugk_mat *Kc, const int *perm, void *s)
{
  const unsigned g = (unsigned)((Kc->n_brows + WG - 1) / WG);
  if (K->bs == 3) k_ssor_permute_vals<3><<<g, WG, 0, S(s)>>>(view(K), K->d_vals, view(Kc), Kc->d_vals, perm);
  else k_ssor_permute_vals<1><<<g, WG, 0, S(s)>>>(view(K), K->d_vals, view(Kc), Kc->d_vals, perm);
  CK_LAUNCH("k_ssor_permute_vals");
  return 0;
}

int ugk_ssor_permute_vec(long n_brows, int bs, const int *perm, const double *src, double *dst, int dir, void *s)
{
  k_ssor_permute_vec<<<vec_blocks(n_brows * bs), WG, 0, S(s)>>>(n_brows, bs, perm, src, dst, dir);
  CK_LAUNCH("k_ssor_permute_vec");
  return 0;
}

// ---- CG with SSOR, Eisenstat's form --------------------------------------
}  // extern "C"

static inline unsigned eis_blocks(long r0, long r1)
{
  return (unsigned)((r1 - 1) / 64 - r0 / 64 + 1);
}

template <int BS, int MODE>
static unsigned eis_sweeps(const ugk_mat *Kc, const int *dpos, int ncol, const long *cb, const double *in1,
                           const double *in2, double *y, double *out1, double *out2, ugk_state *st,
                           double *part, unsigned GB, int gate, hipStream_t s)
{
  const SellView M = view(Kc);
  unsigned pbase = 0;
  for (int k = 0; k < ncol; k++) {
    const int c = (MODE == EIS_BWD_P || MODE == EIS_BWD_X) ? ncol - 1 - k : k;
    const long r0 = cb[c], r1 = cb[c + 1];
    if (r1 <= r0) continue;
    const unsigned g = eis_blocks(r0, r1);
    k_eis_sweep<BS, MODE><<<g, WG, 0, s>>>(M, Kc->d_vals, dpos, r0, r1, r0 / 64, in1, in2, y, out1, out2, st,
                                           MODE == EIS_BWD_P ? part : PART_S(part) + pbase, GB, gate);
    pbase += g;
  }
  return pbase;
}

extern "C" {

int ugk_eis_check_size(const ugk_mat *Kc, int ncol)
{
  if (Kc->n_slices + ncol > 4L * 65536) {
    snprintf(g_err, sizeof(g_err), "SSOR: %ld slices exceed the partial-sum buffer", Kc->n_slices);
    return -1;
  }
  return 0;
}

int ugk_eis_start(const ugk_mat *Kc, const int *dpos, int ncol, const long *cb, long n, const double *r0,
                  double *rh, double *ph, double *xh, double *sv, double *dsq, ugk_state *st, double *part,
                  void *s)
{
  const long nb = Kc->n_brows * Kc->bs;
  if (ugk_eis_check_size(Kc, ncol)) return -1;
  if (Kc->bs == 3) k_eis_dsq<3><<<(unsigned)((nb + WG - 1) / WG), WG, 0, S(s)>>>(view(Kc), Kc->d_vals, dpos, dsq, st);
  else k_eis_dsq<1><<<(unsigned)((nb + WG - 1) / WG), WG, 0, S(s)>>>(view(Kc), Kc->d_vals, dpos, dsq, st);
  k_fill_zero<<<vec_blocks(2 * n), WG, 0, S(s)>>>(2 * n, (unsigned int *)xh);
  unsigned G;
  if (Kc->bs == 3)
    G = eis_sweeps<3, EIS_FWD_R0>(Kc, dpos, ncol, cb, r0, nullptr, sv, rh, ph, st, part, 0, 0, S(s));
  else
    G = eis_sweeps<1, EIS_FWD_R0>(Kc, dpos, ncol, cb, r0, nullptr, sv, rh, ph, st, part, 0, 0, S(s));
  k_eis_sum<<<1, WG, 0, S(s)>>>(PART_S(part), G, PART_B2(part, 0));
  CK_LAUNCH("k_eis_sum");
  return 0;
}

int ugk_eis_iter(const ugk_mat *Kc, const int *dpos, int ncol, const long *cb, long n, double *xh, double *rh,
                 double *ph, double *qh, double *t, double *sv, const double *dsq, const double *x0, double *yx,
                 int check, int rule, ugk_state *st, double *part, void *s)
{
  const unsigned GV = vec_blocks(n);
  unsigned G;
  if (Kc->bs == 3) {
    eis_sweeps<3, EIS_BWD_P>(Kc, dpos, ncol, cb, rh, nullptr, t, ph, nullptr, st, part, GV, 1, S(s));
    G = eis_sweeps<3, EIS_FWD_Q>(Kc, dpos, ncol, cb, ph, t, sv, qh, nullptr, st, part, GV, 1, S(s));
  } else {
    eis_sweeps<1, EIS_BWD_P>(Kc, dpos, ncol, cb, rh, nullptr, t, ph, nullptr, st, part, GV, 1, S(s));
    G = eis_sweeps<1, EIS_FWD_Q>(Kc, dpos, ncol, cb, ph, t, sv, qh, nullptr, st, part, GV, 1, S(s));
  }
  k_eis_upd<<<GV, WG, 0, S(s)>>>(n, xh, rh, ph, qh, st, part, G, GV, check);
  CK_LAUNCH("k_eis_upd");
  if (check) {
    unsigned GX = 0;
    if (rule == 0) {
      if (Kc->bs == 3)
        GX = eis_sweeps<3, EIS_BWD_X>(Kc, dpos, ncol, cb, xh, x0, yx, nullptr, nullptr, st, part, 0, 1, S(s));
      else
        GX = eis_sweeps<1, EIS_BWD_X>(Kc, dpos, ncol, cb, xh, x0, yx, nullptr, nullptr, st, part, 0, 1, S(s));
    }
    if (Kc->bs == 3) k_eis_resid<3><<<GV, WG, 0, S(s)>>>(view(Kc), Kc->d_vals, dpos, rh, dsq, st, part);
    else k_eis_resid<1><<<GV, WG, 0, S(s)>>>(view(Kc), Kc->d_vals, dpos, rh, dsq, st, part);
    k_eis_test<<<1, WG, 0, S(s)>>>(st, part, GV, GX);
    CK_LAUNCH("k_eis_test");
  }
  return 0;
}

int ugk_eis_finish(const ugk_mat *Kc, const int *dpos, int ncol, const long *cb, const double *xh, double *x,
                   double *yx, ugk_state *st, double *part, void *s)
{
  if (Kc->bs == 3)
    eis_sweeps<3, EIS_BWD_X>(Kc, dpos, ncol, cb, xh, x, yx, x, nullptr, st, part, 0, 0, S(s));
  else
    eis_sweeps<1, EIS_BWD_X>(Kc, dpos, ncol, cb, xh, x, yx, x, nullptr, st, part, 0, 0, S(s));
  CK_LAUNCH("k_eis_sweep");
  return 0;
}

int ugk_ssor_iter(const ugk_mat *Kc, int ncol, const long *cb, long n, double *x, double *r, double *zz,
                  double *z, double *p, double *q, const double *dinv, ugk_state *st, double *part, void *s)
{
  const unsigned GV = vec_blocks(n), GA = spmv_blocks(Kc);
  const SellView M = view(Kc);
  int c;
  k_ssor_test<<<1, WG, 0, S(s)>>>(st, part, spmv_blocks(Kc), GV);
  CK_LAUNCH("k_ssor_test");
  for (c = 0; c < ncol; c++) {
    const long r0 = cb[c], r1 = cb[c + 1];
    const unsigned g = (unsigned)(((r1 - 1) / 64 - r0 / 64 + 1 + NW - 1) / NW);
    if (r1 <= r0) continue;
    if (Kc->bs == 3) k_sgs<3, 1><<<g, WG, 0, S(s)>>>(M, Kc->d_vals, r0, r1, r, dinv, zz, st);
    else k_sgs<1, 1><<<g, WG, 0, S(s)>>>(M, Kc->d_vals, r0, r1, r, dinv, zz, st);
  }
  for (c = ncol - 1; c >= 0; c--) {
    const long r0 = cb[c], r1 = cb[c + 1];
    const unsigned g = (unsigned)(((r1 - 1) / 64 - r0 / 64 + 1 + NW - 1) / NW);
    if (r1 <= r0) continue;
    if (Kc->bs == 3) k_sgs<3, 0><<<g, WG, 0, S(s)>>>(M, Kc->d_vals, r0, r1, zz, dinv, z, st);
    else k_sgs<1, 0><<<g, WG, 0, S(s)>>>(M, Kc->d_vals, r0, r1, zz, dinv, z, st);
  }
  CK_LAUNCH("k_sgs");
  k_ssor_dot<<<GV, WG, 0, S(s)>>>(n, r, z, st, part);
  k_ssor_p<<<GV, WG, 0, S(s)>>>(n, z, p, st, part, GV);
  if (Kc->bs == 3) k_pcg_q<3><<<GA, WG, 0, S(s)>>>(M, Kc->d_vals, p, q, st, part);
  else k_pcg_q<1><<<GA, WG, 0, S(s)>>>(M, Kc->d_vals, p, q, st, part);
  k_ssor_upd<<<GV, WG, 0, S(s)>>>(n, x, r, p, q, st, part, GA, nullptr);
  CK_LAUNCH("k_ssor_upd");
  return 0;
}

int ugk_fcg_start(const ugk_mat *K, long n, const double *b, const double *x, double *r, const double *dinv,
                  double *part, void *s)
{
  (void)n;
  if (K->bs == 3) k_fcg_start<3><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, b, x, r, dinv, part);
  else k_fcg_start<1><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, b, x, r, dinv, part);
  CK_LAUNCH("k_fcg_start");
  return 0;
}

int ugk_fcg_head(const ugk_mat *K, long n, const double *r, const double *dinv, unsigned short *x0,
                 ugk_state *st, double *part, void *s)
{
  k_fcg_head<<<vec_blocks(n), WG, 0, S(s)>>>(n, r, dinv, x0, st, part, spmv_blocks(K), vec_blocks(n));
  CK_LAUNCH("k_fcg_head");
  return 0;
}

int ugk_fcg_combine(long n, const double *r, const double *dinv, const float *mask, const float *out,
                    const double *P, const double *Q, int m, const double *zc, double *z, ugk_state *st,
                    double *part, void *s)
{
  if (m < 1 || m > FCG_MAXM) { snprintf(g_err, sizeof(g_err), "FCG: 1 <= m <= %d", FCG_MAXM); return -1; }
  k_fcg_combine<<<vec_blocks(n), WG, 0, S(s)>>>(n, r, dinv, mask, out, P, Q, m, zc, z, st, part);
  CK_LAUNCH("k_fcg_combine");
  return 0;
}

int ugk_fcg_p(long n, const double *r, const double *dinv, const double *z, double *P, int m, ugk_state *st,
              double *part, void *s)
{
  k_fcg_p<<<vec_blocks(n), WG, 0, S(s)>>>(n, r, dinv, z, P, m, st, part, vec_blocks(n));
  CK_LAUNCH("k_fcg_p");
  return 0;
}

int ugk_fcg_q(const ugk_mat *K, long n, const double *P, double *Q, int m, int f32, ugk_state *st,
              double *part, void *s)
{
  if (f32 && !K->d_vals32) { snprintf(g_err, sizeof(g_err), "FCG: no FP32 copy of the matrix"); return -1; }
  if (f32) {
    if (K->bs == 3) k_fcg_q<3, float><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals32, P, Q, n, m, st, part);
    else k_fcg_q<1, float><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals32, P, Q, n, m, st, part);
  } else {
    if (K->bs == 3) k_fcg_q<3, double><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, P, Q, n, m, st, part);
    else k_fcg_q<1, double><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, P, Q, n, m, st, part);
  }
  CK_LAUNCH("k_fcg_q");
  return 0;
}

int ugk_fcg_refresh(const ugk_mat *K, const double *b, const double *x, double *r, const double *dinv,
                    ugk_state *st, double *part, void *s)
{
  if (K->bs == 3) k_fcg_refresh<3><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, b, x, r, dinv, st, part);
  else k_fcg_refresh<1><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, b, x, r, dinv, st, part);
  CK_LAUNCH("k_fcg_refresh");
  return 0;
}

int ugk_fcg_update(const ugk_mat *K, long n, double *x, double *r, const double *P, const double *Q, int m,
                   const double *dinv, ugk_state *st, double *part, void *s)
{
  k_fcg_update<<<vec_blocks(n), WG, 0, S(s)>>>(n, x, r, P, Q, m, dinv, st, part, spmv_blocks(K));
  CK_LAUNCH("k_fcg_update");
  return 0;
}

// FCG(1): the head region holds 6 values per block of the grid that wrote it
static int f1_grid_ok(const ugk_mat *K, long n)
{
  if (spmv_blocks(K) > 65536 || vec_blocks(n) > 65536 || n != K->n_brows * K->bs) {
    snprintf(g_err, sizeof(g_err), "flexible CG: matrix of %ld rows does not fit its partials", n);
    return 0;
  }
  return 1;
}

int ugk_f1_start(const ugk_mat *K, long n, const double *b, const double *x, double *r, const double *dinv,
                 ugk_state *st, double *part, void *s)
{
  if (!f1_grid_ok(K, n)) return -1;
  if (K->bs == 3) k_f1_start<3><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, b, x, r, dinv, st, part);
  else k_f1_start<1><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, b, x, r, dinv, st, part);
  CK_LAUNCH("k_f1_start");
  return 0;
}

int ugk_f1_input0(long n, const double *r, const double *dinv, unsigned short *x0, double *v64, const float *mask,
                  ugk_state *st, double *part, void *s)
{
  k_f1_input0<<<vec_blocks(n), WG, 0, S(s)>>>(n, r, dinv, x0, v64, mask, st, part);
  CK_LAUNCH("k_f1_input0");
  return 0;
}

int ugk_f1_refresh(const ugk_mat *K, long n, const double *b, const double *x, double *r, const double *dinv,
                   const double *p, const double *q, unsigned short *x0, double *v64, const float *mask,
                   ugk_state *st, double *part, void *s)
{
  if (!f1_grid_ok(K, n)) return -1;
  if (K->bs == 3)
    k_f1_refresh<3><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, b, x, r, dinv, p, q, x0, v64, mask, st, part);
  else
    k_f1_refresh<1><<<spmv_blocks(K), WG, 0, S(s)>>>(view(K), K->d_vals, b, x, r, dinv, p, q, x0, v64, mask, st, part);
  CK_LAUNCH("k_f1_refresh");
  return 0;
}

int ugk_f1_combine(long n, const double *r, const double *dinv, const float *mask, const float *out,
                   const double *zc, const double *q, double *z, ugk_state *st, double *part, void *s)
{
  k_f1_combine<<<vec_blocks(n), WG, 0, S(s)>>>(n, r, dinv, mask, out, zc, q, z, st, part);
  CK_LAUNCH("k_f1_combine");
  return 0;
}

int ugk_f1_combine_c0t(long n, const double *r, const double *dinv, const float *mask, const ugk_conv_desc *cd,
                       const double *zc, const double *q, double *z, ugk_state *st, double *part, void *s)
{
  const bool f64 = cd->x_f32 == UGK_ACT_F64;
  if (cd->mode != UGK_CONV_T2 || (cd->x_f32 != UGK_ACT_F32 && !f64) || (cd->cin != 32 && cd->cin != 16) || cd->ldx % 4 ||
      cd->ldx < cd->cin || cd->cout != 3 || cd->cop < 3 || cd->cip < cd->cin || cd->nb != 1 ||
      n != 3L * cd->go.D * cd->go.H * cd->go.W || n / 3 >= (1L << 31) || cd->gi.D != (cd->go.D + 1) / 2 ||
      cd->gi.H != (cd->go.H + 1) / 2 || cd->gi.W != (cd->go.W + 1) / 2) {
    snprintf(g_err, sizeof(g_err), "flexible CG: the network's last layer does not fit the fused combine");
    return -1;
  }
  const long vc = (long)cd->gi.D * cd->gi.H * cd->gi.W;
  const int npb = C0T_NODES;
  if ((vc + npb - 1) / npb > 65536 / 2) {
    snprintf(g_err, sizeof(g_err), "flexible CG: coarse level too large for the fused combine's partials");
    return -1;
  }
  const unsigned G = (unsigned)((vc + npb - 1) / npb);
  if (f64) {
    if (cd->cin == 32) k_f1_combine_c0t<32, double><<<G, C0T_TB, 0, S(s)>>>(n, r, dinv, mask, *cd, zc, q, z, st, part, npb);
    else k_f1_combine_c0t<16, double><<<G, C0T_TB, 0, S(s)>>>(n, r, dinv, mask, *cd, zc, q, z, st, part, npb);
  } else {
    if (cd->cin == 32) k_f1_combine_c0t<32, float><<<G, C0T_TB, 0, S(s)>>>(n, r, dinv, mask, *cd, zc, q, z, st, part, npb);
    else k_f1_combine_c0t<16, float><<<G, C0T_TB, 0, S(s)>>>(n, r, dinv, mask, *cd, zc, q, z, st, part, npb);
  }
  CK_LAUNCH("k_f1_combine_c0t");
  return 0;
}

int ugk_f1_w(const ugk_mat *K, const double *z, double *w, int f32, const ugk_state *st, double *part, void *s)
{
  const SellView M = view(K);
  if (f32) {
    if (!K->d_vals32) {
      snprintf(g_err, sizeof(g_err), "flexible CG: no FP32 copy of the matrix");
      return -1;
    }
    if (K->bs == 3) k_f1_w<3, float><<<spmv_blocks(K), WG, 0, S(s)>>>(M, K->d_vals32, z, w, st, part);
    else k_f1_w<1, float><<<spmv_blocks(K), WG, 0, S(s)>>>(M, K->d_vals32, z, w, st, part);
  } else {
    if (K->bs == 3) k_f1_w<3, double><<<spmv_blocks(K), WG, 0, S(s)>>>(M, K->d_vals, z, w, st, part);
    else k_f1_w<1, double><<<spmv_blocks(K), WG, 0, S(s)>>>(M, K->d_vals, z, w, st, part);
  }
  CK_LAUNCH("k_f1_w");
  return 0;
}

int ugk_f1_update(const ugk_mat *K, long n, double *x, double *r, double *p, double *q, const double *z,
                  const double *w, const double *dinv, unsigned short *x0, double *v64, const float *mask,
                  ugk_state *st, double *part, void *s)
{
  k_f1_update<<<vec_blocks(n), WG, 0, S(s)>>>(n, x, r, p, q, z, w, dinv, x0, v64, mask, st, part, spmv_blocks(K));
  CK_LAUNCH("k_f1_update");
  return 0;
}

static CCGrid ccgrid(const ugk_coarse *C)
{
  CCGrid g;
  g.W = C->W; g.H = C->H; g.D = C->D;
  g.cx = C->cx; g.cy = C->cy; g.cz = C->cz;
  g.Wc = C->Wc; g.Hc = C->Hc; g.Dc = C->Dc;
  g.ncell = C->ncell; g.nc = C->nc; g.ncp = C->ncp;
  return g;
}

int ugk_cc_galerkin(ugk_coarse *C, void *s)
{
  if (C->cx * C->cy * C->cz > UGK_CC_MAX_CELL) {
    snprintf(g_err, sizeof(g_err), "coarse cell of more than %d elements", UGK_CC_MAX_CELL);
    return -1;
  }
  k_cc_galerkin<<<(unsigned)C->ncell, WG, 0, S(s)>>>(ccgrid(C), C->d_G, C->d_E, C->d_nu, C->d_KC);
  CK_LAUNCH("k_cc_galerkin");
  return 0;
}

int ugk_cc_assemble(ugk_coarse *C, void *s)
{
  const long nn = (long)C->Wc * C->Hc * C->Dc;
  k_cc_clear<<<2048, WG, 0, S(s)>>>(C->d_A, C->nc, C->ncp);
  CK_LAUNCH("k_cc_clear");
  k_cc_assemble<<<(unsigned)((nn * 243 + WG - 1) / WG), WG, 0, S(s)>>>(ccgrid(C), C->d_KC, C->d_cm, C->d_A);
  CK_LAUNCH("k_cc_assemble");
  return 0;
}

int ugk_dense_spd_inverse(double *A, long n, double *piv, double *R, double *Cb, void *s)
{
  long K0;
  if (n <= 0 || n % 64) { snprintf(g_err, sizeof(g_err), "dense inverse: n must be a positive multiple of 64"); return -1; }
  if (g_gj_block == 64) {
    for (K0 = 0; K0 < n; K0 += 64) {
      k_gj_pivot<64><<<1, 1024, 0, S(s)>>>(A, n, K0, piv);
      k_gj_panel<64><<<(unsigned)(n / 64), WG, 0, S(s)>>>(A, n, K0, piv, R, Cb);
      k_gj_update<64><<<dim3((unsigned)(n / 64), (unsigned)(n / 64)), WG, 0, S(s)>>>(A, n, K0, piv, R, Cb);
    }
  } else {
    for (K0 = 0; K0 < n; K0 += 32) {
      k_gj_pivot<32><<<1, 1024, 0, S(s)>>>(A, n, K0, piv);
      k_gj_panel<32><<<(unsigned)(n / 64), WG, 0, S(s)>>>(A, n, K0, piv, R, Cb);
      k_gj_update<32><<<dim3((unsigned)(n / 64), (unsigned)(n / 64)), WG, 0, S(s)>>>(A, n, K0, piv, R, Cb);
    }
  }
  CK_LAUNCH("k_gj_update");
  return 0;
}

static int cc_restrict_solve(const ugk_coarse *C, const double *r, const int *stop, void *s)
{
  const CCGrid g = ccgrid(C);
  k_cc_restrict<<<(unsigned)C->ncell, WG, 0, S(s)>>>(g, r, C->d_fm, C->d_cpart, stop);
  k_cc_gather<<<(unsigned)((C->ncp + WG - 1) / WG), WG, 0, S(s)>>>(g, C->d_cpart, C->d_cm, C->d_g, stop);
  k_cc_gemv<<<(unsigned)((C->ncp + NW - 1) / NW), WG, 0, S(s)>>>(C->d_A, C->ncp, C->d_g, C->d_y, stop);
  CK_LAUNCH("k_cc_gemv");
  return 0;
}

int ugk_cc_iter(const ugk_mat *K, const ugk_coarse *C, long n, double *x, double *r, double *z, double *p,
                double *q, const double *dinv, ugk_state *st, double *part, void *s)
{
  const unsigned GV = vec_blocks(n), GA = spmv_blocks(K);
  const SellView M = view(K);
  if (K->bs != 3 || n != 3L * C->W * C->H * C->D) {
    snprintf(g_err, sizeof(g_err), "two-level: a 3x3-block matrix on the coarse grid's nodes is needed");
    return -1;
  }
  k_ssor_test<<<1, WG, 0, S(s)>>>(st, part, GA, GV);
  CK_LAUNCH("k_ssor_test");
  if (cc_restrict_solve(C, r, &st->done, s)) return -1;
  k_cc_z<1><<<GV, WG, 0, S(s)>>>(ccgrid(C), C->d_y, C->d_fm, r, dinv, z, &st->done, part);
  k_ssor_p<<<GV, WG, 0, S(s)>>>(n, z, p, st, part, GV);
  k_pcg_q<3><<<GA, WG, 0, S(s)>>>(M, K->d_vals, p, q, st, part);
  k_ssor_upd<<<GV, WG, 0, S(s)>>>(n, x, r, p, q, st, part, GA, nullptr);
  CK_LAUNCH("k_ssor_upd");
  return 0;
}

int ugk_cc_iter_mixed(const ugk_mat *K, const ugk_coarse *C, long n, const double *b, double *x, double *r,
                      double *z, double *p, double *q, const double *dinv, ugk_state *st, double *part, void *s)
{
  const unsigned GV = vec_blocks(n), GA = spmv_blocks(K);
  const SellView M = view(K);
  if (K->bs != 3 || n != 3L * C->W * C->H * C->D || !K->d_vals32) {
    snprintf(g_err, sizeof(g_err), "two-level mixed: a 3x3-block matrix with its FP32 copy on the grid's nodes");
    return -1;
  }
  k_fcg_refresh<3><<<GA, WG, 0, S(s)>>>(M, K->d_vals, b, x, r, dinv, st, part);
  k_fcg_head<<<GV, WG, 0, S(s)>>>(n, r, dinv, NULL, st, part, GA, GV);
  CK_LAUNCH("k_fcg_head");
  if (cc_restrict_solve(C, r, &st->done, s)) return -1;
  k_cc_z<1><<<GV, WG, 0, S(s)>>>(ccgrid(C), C->d_y, C->d_fm, r, dinv, z, &st->done, part);
  k_ssor_p<<<GV, WG, 0, S(s)>>>(n, z, p, st, part, GV);
  k_pcg_q<3, float><<<GA, WG, 0, S(s)>>>(M, K->d_vals32, p, q, st, part);
  k_ssor_upd<<<GV, WG, 0, S(s)>>>(n, x, r, p, q, st, part, GA, nullptr);
  CK_LAUNCH("k_ssor_upd");
  return 0;
}

int ugk_mx64_iter(const ugk_mat *K, long n, const double *b, double *x, double *r, double *p, double *q,
                  const double *dinv, ugk_state *st, double *part, int refresh, void *s)
{
  const unsigned GV = vec_blocks(n), GA = spmv_blocks(K);
  const SellView M = view(K);
  if (!K->d_vals32) {
    snprintf(g_err, sizeof(g_err), "mixed PCG: no FP32 copy of the matrix");
    return -1;
  }
  if (refresh) {
    if (K->bs == 3) k_fcg_refresh<3, true><<<GA, WG, 0, S(s)>>>(M, K->d_vals, b, x, r, dinv, st, part);
    else k_fcg_refresh<1, true><<<GA, WG, 0, S(s)>>>(M, K->d_vals, b, x, r, dinv, st, part);
    CK_LAUNCH("k_fcg_refresh");
  }
  k_mx64_p<<<GV, WG, 0, S(s)>>>(n, r, dinv, p, st, part, GA, GV);
  if (K->bs == 3) k_pcg_q<3, float><<<GA, WG, 0, S(s)>>>(M, K->d_vals32, p, q, st, part);
  else k_pcg_q<1, float><<<GA, WG, 0, S(s)>>>(M, K->d_vals32, p, q, st, part);
  k_ssor_upd<<<GV, WG, 0, S(s)>>>(n, x, r, p, q, st, part, GA, dinv);
  CK_LAUNCH("k_ssor_upd");
  return 0;
}

int ugk_cc_apply(const ugk_coarse *C, long n, const double *r, double *zc, const ugk_state *st, void *s)
{
  if (n != 3L * C->W * C->H * C->D) { snprintf(g_err, sizeof(g_err), "two-level: size mismatch"); return -1; }
  if (cc_restrict_solve(C, r, &st->done, s)) return -1;
  k_cc_z<0><<<vec_blocks(n), WG, 0, S(s)>>>(ccgrid(C), C->d_y, C->d_fm, r, NULL, zc, &st->done, NULL);
  CK_LAUNCH("k_cc_z");
  return 0;
}

}  // extern "C"
