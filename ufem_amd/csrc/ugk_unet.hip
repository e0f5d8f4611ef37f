// ugk_unet.hip -- BF16 matrix-core kernels of the learned preconditioner's
// 3-D convolution network (gfx950).
//
// All activations are bf16, channels-last on the structured node grid, so a
// (voxel, 4 channels) group is one 8-byte load.  A 3x3x3 convolution is an
// implicit GEMM on the 16x16x16 bf16 MFMA: out-channels x voxels, summed over
// 27 taps x in-channels.  The same kernel computes data gradients: a
// stride-1 layer's gradient is the convolution with the flipped, transposed
// taps; a stride-2 layer's is the "T2" transpose mode, where output voxel f
// takes input (f - off) / 2 when that is integral.  Weight gradients stage a
// 64-voxel chunk of the output gradient and of the gathered input, both
// transposed, in LDS, and each block sums one tap over a slice of voxels;
// slices are summed in a fixed order afterwards (no atomics, so training
// steps repeat bit for bit).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "ugk.h"

int ugk_hip_fail(const char *what, hipError_t e);   // ugk_kernels.hip

#define CK(call)                                            \
  do {                                                      \
    hipError_t e_ = (call);                                 \
    if (e_ != hipSuccess) return ugk_hip_fail(#call, e_);   \
  } while (0)
#define CK_LAUNCH(what) CK(hipGetLastError())

typedef unsigned short bf16;
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

#define SLOPE 0.1f

__device__ __forceinline__ float bf2f(bf16 h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ float bf2f(short h) { return __uint_as_float((unsigned)(unsigned short)h << 16); }
// round to nearest even
__device__ __forceinline__ bf16 f2bf(float f)
{
  unsigned u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16)0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}

static inline hipStream_t S(void *s) { return (hipStream_t)s; }
__host__ __device__ static inline long vox(ugk_grid3 g) { return (long)g.D * g.H * g.W; }

// input voxel of tap (dz, dy, dx) for output voxel (z, y, x); -1 if none
template <int MODE>
__device__ __forceinline__ long tap_src(int z, int y, int x, int dz, int dy, int dx, const ugk_grid3 &gi)
{
  int zi, yi, xi;
  if (MODE == UGK_CONV_S1) {
    zi = z + dz; yi = y + dy; xi = x + dx;
  } else if (MODE == UGK_CONV_S2) {
    zi = 2 * z + dz; yi = 2 * y + dy; xi = 2 * x + dx;
  } else {
    const int a = z - dz, b = y - dy, c = x - dx;
    if ((a | b | c) & 1) return -1;        // odd (also catches -1)
    zi = a >> 1; yi = b >> 1; xi = c >> 1;
  }
  if (zi < 0 || zi >= gi.D || yi < 0 || yi >= gi.H || xi < 0 || xi >= gi.W) return -1;
  return ((long)zi * gi.H + yi) * gi.W + xi;
}

// --------------------------------------------------------------------------
// convolution: each wave computes NCO x 16 output channels for NV x 16
// voxels (A = weights [co][ci], B = activations [ci][voxel]; a lane of the
// accumulator holds 4 consecutive channels of one voxel -> 8/16-byte stores).
// XF = 1: f32 activations, each split exactly into three bf16 parts
// (x = hi + mid + lo to 24 bits) that are multiplied by the same bf16
// weights, so the product is the f32 activation times the bf16 weight.
// XF = 2: the same three parts, stored split by the producer (no per-load
// split work; the products are bit for bit those of XF = 1).
// --------------------------------------------------------------------------
// B fragment of 4 channels: f32 loaded raw (split after the load arrives),
// the three stored parts, or bf16 as is
struct Split4 { bf16x4 p[3]; };
template <int XF> struct Frag;
template <> struct Frag<2> { typedef Split4 T; };
template <> struct Frag<1> { typedef f32x4 T; };
template <> struct Frag<0> { typedef bf16x4 T; };

// off: element offset v * ld + c of channel c (c % 4 == 0) of voxel v
template <int XF>
__device__ __forceinline__ typename Frag<XF>::T load_raw(const void *x, long off)
{
  if constexpr (XF == 2) {
    const bf16 *q = (const bf16 *)x + off * 3;
    Split4 r;
    r.p[0] = *(const bf16x4 *)q;
    r.p[1] = *(const bf16x4 *)(q + 4);
    r.p[2] = *(const bf16x4 *)(q + 8);
    return r;
  } else if constexpr (XF == 1) {
    return *(const f32x4 *)((const float *)x + off);
  } else {
    return *(const bf16x4 *)((const bf16 *)x + off);
  }
}

// exact three-way bf16 split of 4 f32 values by truncation: hi keeps the
// top 8 significant bits, mid the next 8 of the remainder, lo the last 8
// (x = hi + mid + lo exactly, the subtractions are exact); invalid lanes
// give zeros
__device__ __forceinline__ void split3(const f32x4 &x, bool ok, bf16x4 (&b)[3])
{
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const unsigned u = ok ? __float_as_uint(x[j]) : 0u;
    const unsigned h = u & 0xffff0000u;
    const float r1 = __uint_as_float(u) - __uint_as_float(h);
    const unsigned m = __float_as_uint(r1) & 0xffff0000u;
    const float r2 = r1 - __uint_as_float(m);
    b[0][j] = (short)(h >> 16);
    b[1][j] = (short)(m >> 16);
    b[2][j] = (short)(__float_as_uint(r2) >> 16);
  }
}

// epilogue operands (emul, ga, gb) of 4 consecutive channels co.. of output
// voxel g: 16-byte loads where the layout allows, issued by the callers
// ahead of their products so that the latency overlaps
struct Epi { f32x4 em, ga, gb; };

__device__ __forceinline__ bool al16(const float *p, int ld) { return !(((unsigned long)p & 15) | (ld & 3)); }

__device__ __forceinline__ Epi epi_load(const ugk_conv_desc &d, long g, int co, bool valid)
{
  Epi e;
  e.em = (f32x4){1.f, 1.f, 1.f, 1.f};
  e.ga = (f32x4){0.f, 0.f, 0.f, 0.f};
  e.gb = e.ga;
  const int nj = d.cout - co < 4 ? d.cout - co : 4;
  if (!valid || nj <= 0) return e;
  if (d.emul) {
    if (nj == 4 && al16(d.emul, d.ldm)) e.em = *(const f32x4 *)(d.emul + g * d.ldm + co);
    else for (int j = 0; j < nj; j++) e.em[j] = d.emul[g * d.ldm + co + j];
  }
  if (d.ga) {
    if (nj == 4 && al16(d.ga, d.ldga) && al16(d.gb, d.ldgb)) {
      e.ga = *(const f32x4 *)(d.ga + g * d.ldga + co);
      e.gb = *(const f32x4 *)(d.gb + g * d.ldgb + co);
    } else {
      for (int j = 0; j < nj; j++) {
        e.ga[j] = d.ga[g * d.ldga + co + j];
        e.gb[j] = d.gb[g * d.ldgb + co + j];
      }
    }
  }
  return e;
}

// epilogue with operands loaded by epi_load: v = acc (* emul) (+ ga gb)
__device__ __forceinline__ void store_pre(const ugk_conv_desc &d, long g, int co, const f32x4 &acc, const Epi &e)
{
  const int nj = d.cout - co < 4 ? d.cout - co : 4;
  float v[4];
#pragma unroll
  for (int j = 0; j < 4; j++) v[j] = acc[j];
  if (d.emul)
    for (int j = 0; j < nj; j++) v[j] *= e.em[j];
  if (d.ga)
    for (int j = 0; j < nj; j++) v[j] += e.ga[j] * e.gb[j];
  if (d.y_f32) {
    float *y = (float *)d.y + g * d.ldy + co;
    if (nj == 4) *(f32x4 *)y = (f32x4){v[0], v[1], v[2], v[3]};
    else for (int j = 0; j < nj; j++) y[j] = v[j];
  } else {
    bf16 *y = (bf16 *)d.y + g * d.ldy + co;
    if (nj == 4) *(bf16x4 *)y = (bf16x4){(short)f2bf(v[0]), (short)f2bf(v[1]), (short)f2bf(v[2]), (short)f2bf(v[3])};
    else for (int j = 0; j < nj; j++) y[j] = f2bf(v[j]);
  }
}

__device__ __forceinline__ void store_out(const ugk_conv_desc &d, long g, int co, const f32x4 &acc)
{
  store_pre(d, g, co, acc, epi_load(d, g, co, true));
}

template <int XF>
__device__ __forceinline__ void mfma_frag(const bf16x4 &a, const typename Frag<XF>::T &raw, bool ok, f32x4 &acc)
{
  if constexpr (XF == 2) {
    const bf16x4 z = (bf16x4){0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, ok ? raw.p[2] : z, acc, 0, 0, 0);   // smallest part first
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, ok ? raw.p[1] : z, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, ok ? raw.p[0] : z, acc, 0, 0, 0);
  } else if constexpr (XF == 1) {
    bf16x4 bx[3];
    split3(raw, ok, bx);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bx[2], acc, 0, 0, 0);   // smallest part first
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bx[1], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bx[0], acc, 0, 0, 0);
  } else {
    const bf16x4 bx = ok ? raw : (bf16x4){0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bx, acc, 0, 0, 0);
  }
}

// the bf16 parts of a B fragment (zeros when !ok), and the products of one
// weight fragment with them in the order of mfma_frag (smallest part
// first), so a fragment split once serves several weight fragments with the
// same results
template <int XF>
__device__ __forceinline__ void to_parts(const typename Frag<XF>::T &raw, bool ok, bf16x4 (&b)[3])
{
  const bf16x4 z = (bf16x4){0, 0, 0, 0};
  if constexpr (XF == 2) {
    b[0] = ok ? raw.p[0] : z;
    b[1] = ok ? raw.p[1] : z;
    b[2] = ok ? raw.p[2] : z;
  } else if constexpr (XF == 1) {
    split3(raw, ok, b);
  } else {
    b[0] = ok ? raw : z;
  }
}
template <int XF>
__device__ __forceinline__ void mfma_parts(const bf16x4 &a, const bf16x4 (&b)[3], f32x4 &acc)
{
  if constexpr (XF == 0) {
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[0], acc, 0, 0, 0);
  } else {
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[2], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[1], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[0], acc, 0, 0, 0);
  }
}

// Stride-1 and stride-2 convolution.  Each wave computes one 16-channel
// output group (blockIdx.y) for NV x 16 voxels; the 27 taps are unrolled
// without branches (clamped addresses, zeroed fragments), so the gathers of
// many taps are in flight together.  PK: the input has 4 channels, and the
// four lane groups of the MFMA's K dimension take four taps at once.
// KS = 1: the (16-channel chunk, dz plane) units of the sum in a rolled
// loop (fewer registers, more waves resident); KS >= 4: the KS waves of a
// block (64 KS threads) share one tile and take every KS-th unit, which the
// block adds in wave order in LDS -- KS times the waves and 1 / KS of the
// chain (on the small coarse grids KS is the number of units, so that every
// gather of the tile is in flight at once).  KS = 0: all three planes of a
// chunk unrolled.  NG (KS >= 4, NV = 1): the wave computes NG 16-channel
// output groups (blockIdx.y NG .. NG + NG - 1) from one set of gathered and
// split input fragments.
template <int NV, int MODE, int XF, int PK, int KS, int NG = 1>
__global__ __launch_bounds__(KS >= 4 ? 64 * KS : 256) void k_conv(ugk_conv_desc d)
{
  static_assert(NG == 1 || (KS >= 4 && NV == 1), "output groups per wave need the KS >= 4 form");
  if (d.skip && *d.skip) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, kq = lane >> 4;
  const int mg = blockIdx.y * NG;
  const long Vo = vox(d.go), Vi = vox(d.gi);
  const long total = Vo * d.nb;
  const long base = (KS >= 4 ? (long)blockIdx.x : ((long)blockIdx.x * 4 + wave)) * (16 * NV);
  if (base >= total) return;               // uniform over the block when KS >= 4
  const unsigned HWo = (unsigned)(d.go.H * d.go.W), Wo = (unsigned)d.go.W;
  int vz[NV], vy[NV], vx[NV];
  long vb[NV];
  bool vok[NV];
#pragma unroll
  for (int n = 0; n < NV; n++) {
    // 32-bit index arithmetic (ugk_conv checks that the sizes fit)
    const long g = base + n * 16 + col;
    vok[n] = g < total;
    const unsigned gg = vok[n] ? (unsigned)g : 0u;
    const unsigned b = gg / (unsigned)Vo;
    const unsigned v = gg - b * (unsigned)Vo;
    const unsigned z = v / HWo, r = v - z * HWo, y = r / Wo;
    vz[n] = (int)z;
    vy[n] = (int)y;
    vx[n] = (int)(r - y * Wo);
    vb[n] = (long)b * Vi;
  }
  f32x4 acc[NV];
#pragma unroll
  for (int n = 0; n < NV; n++) acc[n] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const bf16 *wrow = d.w + (long)(mg * 16 + col) * d.cip;
  const long wstep = (long)d.cop * d.cip;

  Epi ep[NV];
  if constexpr (KS >= 4 && NG > 1) {
    // one tile, NG output groups: every fragment split once for all groups
    f32x4 ag[NG];
    Epi eg[NG];
#pragma unroll
    for (int g = 0; g < NG; g++) {
      ag[g] = (f32x4){0.f, 0.f, 0.f, 0.f};
      const int co = (mg + g) * 16 + 4 * kq;
      eg[g] = epi_load(d, base + col, co, wave == 0 && vok[0] && co < d.cout);
    }
    const int U = (d.cin + 15) / 16 * 3;
    for (int u = wave; u < U; u += KS) {
      const int c = (u / 3) * 16 + 4 * kq, pz = u % 3;
      const bool cok = c < d.cin;
      const int cs = cok ? c : 0;
      typename Frag<XF>::T raw[9];
      bool ok[9];
      bf16x4 a[NG][9];
#pragma unroll
      for (int k = 0; k < 9; k++) {
        const int t = 9 * pz + k;
        const int dz = pz - 1, dy = k / 3 - 1, dx = k % 3 - 1;
        const long sidx = tap_src<MODE>(vz[0], vy[0], vx[0], dz, dy, dx, d.gi);
        ok[k] = vok[0] && cok && sidx >= 0;
        raw[k] = load_raw<XF>(d.x, (ok[k] ? vb[0] + sidx : 0) * d.ldx + cs);
#pragma unroll
        for (int g = 0; g < NG; g++) a[g][k] = *(const bf16x4 *)(wrow + (long)g * 16 * d.cip + (d.flip ? 26 - t : t) * wstep + c);
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int k = 0; k < 9; k++) {
        bf16x4 bx[3];
        to_parts<XF>(raw[k], ok[k], bx);
#pragma unroll
        for (int g = 0; g < NG; g++) mfma_parts<XF>(a[g][k], bx, ag[g]);
      }
    }
    __shared__ f32x4 redg[KS][NG][64];
#pragma unroll
    for (int g = 0; g < NG; g++) redg[wave][g][lane] = ag[g];
    __syncthreads();
    if (wave || !vok[0]) return;
#pragma unroll
    for (int g = 0; g < NG; g++) {
      f32x4 a = redg[0][g][lane];
#pragma unroll
      for (int w = 1; w < KS; w++) a += redg[w][g][lane];
      const int co = (mg + g) * 16 + 4 * kq;
      if (co < d.cout) store_pre(d, base + col, co, a, eg[g]);
    }
    return;
  }
  if constexpr (KS >= 4) {
    const int co = mg * 16 + 4 * kq;
#pragma unroll
    for (int n = 0; n < NV; n++) ep[n] = epi_load(d, base + n * 16 + col, co, wave == 0 && vok[n] && co < d.cout);
  }
  if constexpr (KS >= 1) {
    const int U = (d.cin + 15) / 16 * 3;
    for (int u = KS >= 4 ? wave : 0; u < U; u += KS) {
      const int c = (u / 3) * 16 + 4 * kq, pz = u % 3;
      const bool cok = c < d.cin;
      const int cs = cok ? c : 0;
      typename Frag<XF>::T raw[9][NV];
      bool ok[9][NV];
      bf16x4 a[9];
#pragma unroll
      for (int k = 0; k < 9; k++) {
        const int t = 9 * pz + k;
        const int dz = pz - 1, dy = k / 3 - 1, dx = k % 3 - 1;
#pragma unroll
        for (int n = 0; n < NV; n++) {
          const long sidx = tap_src<MODE>(vz[n], vy[n], vx[n], dz, dy, dx, d.gi);
          ok[k][n] = vok[n] && cok && sidx >= 0;
          raw[k][n] = load_raw<XF>(d.x, (ok[k][n] ? vb[n] + sidx : 0) * d.ldx + cs);
        }
        a[k] = *(const bf16x4 *)(wrow + (d.flip ? 26 - t : t) * wstep + c);
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int k = 0; k < 9; k++)
#pragma unroll
        for (int n = 0; n < NV; n++) mfma_frag<XF>(a[k], raw[k][n], ok[k][n], acc[n]);
    }
    if constexpr (KS >= 4) {
      __shared__ f32x4 red[KS][NV][64];
#pragma unroll
      for (int n = 0; n < NV; n++) red[wave][n][lane] = acc[n];
      __syncthreads();
      if (wave) return;
#pragma unroll
      for (int n = 0; n < NV; n++) {
        f32x4 a = red[0][n][lane];
#pragma unroll
        for (int w = 1; w < KS; w++) a += red[w][n][lane];
        acc[n] = a;
      }
    }
  } else if constexpr (PK) {
    // tap groups of 4: lane group kq takes tap 4 q + kq, channels 0..3; all
    // 7 groups' fragments are loaded before the first product
    typename Frag<XF>::T raw[7][NV];
    bool ok[7][NV];
    bf16x4 a[7];
#pragma unroll
    for (int q = 0; q < 7; q++) {
      const int t = 4 * q + kq;
      const bool tok = t < 27;
      const int tt = tok ? t : 0;
      const int dz = tt / 9 - 1, dy = (tt / 3) % 3 - 1, dx = tt % 3 - 1;
#pragma unroll
      for (int n = 0; n < NV; n++) {
        const long sidx = tap_src<MODE>(vz[n], vy[n], vx[n], dz, dy, dx, d.gi);
        ok[q][n] = vok[n] && tok && sidx >= 0;
        raw[q][n] = load_raw<XF>(d.x, (ok[q][n] ? vb[n] + sidx : 0) * d.ldx);
      }
      a[q] = *(const bf16x4 *)(wrow + (d.flip ? 26 - tt : tt) * wstep);
      if (!tok) a[q] = (bf16x4){0, 0, 0, 0};
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 0; q < 7; q++)
#pragma unroll
      for (int n = 0; n < NV; n++) mfma_frag<XF>(a[q], raw[q][n], ok[q][n], acc[n]);
  } else {
    for (int c0 = 0; c0 < d.cin; c0 += 16) {
      const int c = c0 + 4 * kq;
      const bool cok = c < d.cin;
      const int cs = cok ? c : 0;              // in-bounds channel for masked lanes
      // one dz plane (9 taps) of gathers in flight before its products
#pragma unroll
      for (int pz = 0; pz < 3; pz++) {
        typename Frag<XF>::T raw[9][NV];
        bool ok[9][NV];
        bf16x4 a[9];
#pragma unroll
        for (int k = 0; k < 9; k++) {
          const int t = 9 * pz + k;
          const int dz = pz - 1, dy = k / 3 - 1, dx = k % 3 - 1;
#pragma unroll
          for (int n = 0; n < NV; n++) {
            const long sidx = tap_src<MODE>(vz[n], vy[n], vx[n], dz, dy, dx, d.gi);
            ok[k][n] = vok[n] && cok && sidx >= 0;
            raw[k][n] = load_raw<XF>(d.x, (ok[k][n] ? vb[n] + sidx : 0) * d.ldx + cs);
          }
          a[k] = *(const bf16x4 *)(wrow + (d.flip ? 26 - t : t) * wstep + c);
        }
        __builtin_amdgcn_sched_barrier(0);   // keep the plane's gathers ahead of its products
#pragma unroll
        for (int k = 0; k < 9; k++)
#pragma unroll
          for (int n = 0; n < NV; n++) mfma_frag<XF>(a[k], raw[k][n], ok[k][n], acc[n]);
      }
    }
  }
  // lane holds channels mg*16 + 4 kq + j of voxel base + n*16 + col
  const int co = mg * 16 + 4 * kq;
  if (co >= d.cout) return;
#pragma unroll
  for (int n = 0; n < NV; n++) {
    if (!vok[n]) continue;
    if constexpr (KS >= 4) store_pre(d, base + n * 16 + col, co, acc[n], ep[n]);
    else store_out(d, base + n * 16 + col, co, acc[n]);
  }
}

// Stride-2 transpose (coarse input gi -> fine output go): y[f] = sum_t W[t]
// x[(f - off_t) / 2] over the taps with f - off_t even.  Output voxels are
// taken parity class by parity class (f = 2c + p), so every lane of a wave
// has the same 1, 2, 4 or 8 live taps: per axis, p = 0 takes off 0 at c, and
// p = 1 takes off +1 at c and off -1 at c + 1.  Work items are
// [nb][8][Vc rounded up to 16] (c over the coarse grid, rows of 16 never
// straddle two classes).  The class is a compile-time case, so the gathers
// of all live taps of a 16-channel chunk are in flight before its products.
template <int PZ, int PY, int PX, int XF>
__device__ __forceinline__ void t2_class(const ugk_conv_desc &d, const bf16 *wrow, long wstep, long xb, int cz,
                                         int cy, int cx, bool fok, int kq, f32x4 &acc, int cfirst, int cstep)
{
  constexpr int NY = PY + 1, NX = PX + 1, NT = (PZ + 1) * NY * NX;
  for (int c0 = cfirst; c0 < d.cin; c0 += cstep) {
    const int c = c0 + 4 * kq;
    const bool cok = c < d.cin;
    const int cs = cok ? c : 0;
    typename Frag<XF>::T raw[NT];
    bool ok[NT];
    bf16x4 a[NT];
#pragma unroll
    for (int j = 0; j < NT; j++) {
      const int iz = j / (NY * NX), iy = (j / NX) % NY, ix = j % NX;
      // offset +1 reads c, offset -1 reads c + 1; parity 0 has offset 0
      const int dz = PZ ? 1 - 2 * iz : 0, dy = PY ? 1 - 2 * iy : 0, dx = PX ? 1 - 2 * ix : 0;
      const int sz = cz + iz, sy = cy + iy, sx = cx + ix;
      ok[j] = fok && cok && sz < d.gi.D && sy < d.gi.H && sx < d.gi.W;
      const long src = ok[j] ? xb + ((long)sz * d.gi.H + sy) * d.gi.W + sx : 0;
      raw[j] = load_raw<XF>(d.x, src * d.ldx + cs);
      const int t = 9 * (dz + 1) + 3 * (dy + 1) + (dx + 1);
      a[j] = *(const bf16x4 *)(wrow + t * wstep + c);
    }
    __builtin_amdgcn_sched_barrier(0);   // the chunk's gathers ahead of its products
#pragma unroll
    for (int j = 0; j < NT; j++) mfma_frag<XF>(a[j], raw[j], ok[j], acc);
  }
}

// KS = 4: the four waves of a block share one group of 16 items and take
// every fourth 16-channel chunk of the input, added in wave order in LDS
// (four times the waves on the small coarse grids); KS = 0: one group of
// items per wave, all chunks in its own chain.
template <int XF, int KS>
__global__ __launch_bounds__(256) void k_conv_t2(ugk_conv_desc d)
{
  if (d.skip && *d.skip) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, kq = lane >> 4;
  const int mg = blockIdx.y;
  const long Vc = vox(d.gi), Vf = vox(d.go);
  const long Vp = (Vc + 15) / 16 * 16;
  const long total = Vp * 8 * d.nb;
  const long base = (KS ? (long)blockIdx.x : ((long)blockIdx.x * 4 + wave)) * 16;
  if (base >= total) return;               // uniform over the block when KS = 4
  // 32-bit index arithmetic (ugk_conv checks that the sizes fit)
  const unsigned cls = (unsigned)base / (unsigned)Vp;   // (b, p), uniform over the wave
  const int p = (int)(cls & 7);
  const long b = cls >> 3;
  const unsigned ci = (unsigned)base - cls * (unsigned)Vp + col;   // coarse voxel of this lane
  const bool cin_ok = ci < (unsigned)Vc;
  const unsigned HWc = (unsigned)(d.gi.H * d.gi.W);
  const int cz = cin_ok ? (int)(ci / HWc) : 0;
  const int rr = cin_ok ? (int)(ci - (unsigned)cz * HWc) : 0;
  const int cy = rr / d.gi.W, cx = rr - cy * d.gi.W;
  const int pz = (p >> 2) & 1, py = (p >> 1) & 1, px = p & 1;
  const int fz = 2 * cz + pz, fy = 2 * cy + py, fx = 2 * cx + px;
  const bool fok = cin_ok && fz < d.go.D && fy < d.go.H && fx < d.go.W;
  f32x4 acc = (f32x4){0.f, 0.f, 0.f, 0.f};
  const bf16 *wrow = d.w + (long)(mg * 16 + col) * d.cip;
  const long wstep = (long)d.cop * d.cip;
  const long xb = b * Vc;
  const int co = mg * 16 + 4 * kq;
  const long gout = b * Vf + ((long)fz * d.go.H + fy) * d.go.W + fx;
  const int c1 = KS ? 16 * wave : 0, cs = KS ? 16 * KS : 16;
  const Epi ep = epi_load(d, gout, co, (!KS || wave == 0) && fok && co < d.cout);
  switch (p) {                              // wave-uniform
    case 0: t2_class<0, 0, 0, XF>(d, wrow, wstep, xb, cz, cy, cx, fok, kq, acc, c1, cs); break;
    case 1: t2_class<0, 0, 1, XF>(d, wrow, wstep, xb, cz, cy, cx, fok, kq, acc, c1, cs); break;
    case 2: t2_class<0, 1, 0, XF>(d, wrow, wstep, xb, cz, cy, cx, fok, kq, acc, c1, cs); break;
    case 3: t2_class<0, 1, 1, XF>(d, wrow, wstep, xb, cz, cy, cx, fok, kq, acc, c1, cs); break;
    case 4: t2_class<1, 0, 0, XF>(d, wrow, wstep, xb, cz, cy, cx, fok, kq, acc, c1, cs); break;
    case 5: t2_class<1, 0, 1, XF>(d, wrow, wstep, xb, cz, cy, cx, fok, kq, acc, c1, cs); break;
    case 6: t2_class<1, 1, 0, XF>(d, wrow, wstep, xb, cz, cy, cx, fok, kq, acc, c1, cs); break;
    default: t2_class<1, 1, 1, XF>(d, wrow, wstep, xb, cz, cy, cx, fok, kq, acc, c1, cs); break;
  }
  if constexpr (KS) {
    __shared__ f32x4 red[KS][64];
    red[wave][lane] = acc;
    __syncthreads();
    if (wave) return;
    f32x4 a = red[0][lane];
#pragma unroll
    for (int w = 1; w < KS; w++) a += red[w][lane];
    acc = a;
  }
  if (co >= d.cout || !fok) return;
  store_pre(d, gout, co, acc, ep);
}

// Stride-1 convolution to at most 4 output channels (the network's last
// layer) on the vector ALUs in f32.  Four lanes share one output voxel, each
// taking one 4-channel quarter of every 16 input channels, so a lane has
// the 9 tap gathers of a dz plane in flight at once; the four partial sums are
// added by lane shuffles in a fixed order.  The bf16 weights of the output
// channels are widened to f32 in LDS.  A 16-row matrix-core tile would
// leave 12 of its rows idle here.
#define SO_MAX_CIN 64
template <int C16, int NCO>
__global__ __launch_bounds__(256) void k_conv_small_out(ugk_conv_desc d)
{
  __shared__ f32x4 sw[27][4][4 * C16];          // [tap][co][ci / 4]
  if (d.skip && *d.skip) return;
  for (int i = threadIdx.x; i < 27 * 4 * 16 * C16; i += 256) {
    const int t = i / (64 * C16), r = i - t * 64 * C16, co = r / (16 * C16), ci = r - co * 16 * C16;
    const int tw = d.flip ? 26 - t : t;
    const bool in = co < d.cout && ci < d.cin;
    ((float *)&sw[0][0][0])[i] = in ? bf2f(d.w[((long)tw * d.cop + co) * d.cip + ci]) : 0.f;
  }
  __syncthreads();
  const long Vo = vox(d.go), Vi = vox(d.gi);
  const long total = Vo * d.nb;
  const int HWo = d.go.H * d.go.W;
  const int cq = threadIdx.x & 3;
  // the loop condition depends on the voxel only, so the four lanes of a
  // voxel stay together for the shuffles
  for (long g = ((long)blockIdx.x * 256 + threadIdx.x) >> 2; g < total; g += (long)gridDim.x * 64) {
    const long b = g / Vo;
    const int v = (int)(g - b * Vo);
    const int z = v / HWo, r = v - z * HWo, y = r / d.go.W, x = r - y * d.go.W;
    f32x4 acc = (f32x4){0.f, 0.f, 0.f, 0.f};
    const float *xb = (const float *)d.x + b * Vi * d.ldx;
#pragma unroll 1
    for (int jz = 0; jz < 3 * C16; jz++) {       // (16-channel group, dz plane)
      const int j = jz / 3, dz = jz - 3 * j - 1;
      const int c = 16 * j + 4 * cq;
      f32x4 xv[9];
#pragma unroll
      for (int t = 0; t < 9; t++) {
        const long src = tap_src<UGK_CONV_S1>(z, y, x, dz, t / 3 - 1, t % 3 - 1, d.gi);
        const f32x4 v4 = *(const f32x4 *)(xb + (src < 0 ? 0 : src) * d.ldx + c);
        xv[t] = src < 0 ? (f32x4){0.f, 0.f, 0.f, 0.f} : v4;
      }
#pragma unroll
      for (int t = 0; t < 9; t++)
#pragma unroll
        for (int co = 0; co < NCO; co++) {
          const f32x4 w4 = sw[9 * (dz + 1) + t][co][4 * j + cq];
          acc[co] += w4[0] * xv[t][0] + w4[1] * xv[t][1] + w4[2] * xv[t][2] + w4[3] * xv[t][3];
        }
    }
#pragma unroll
    for (int co = 0; co < NCO; co++) {
      acc[co] += __shfl_xor(acc[co], 1);
      acc[co] += __shfl_xor(acc[co], 2);
    }
    if (cq == 0) store_out(d, g, 0, acc);
  }
}

template <int NV, int XF, int PK, int KS = 0, int NG = 1>
static void launch_conv_nv(const ugk_conv_desc &d, dim3 grid, hipStream_t s)
{
  const int threads = KS >= 4 ? 64 * KS : 256;
  if (d.mode == UGK_CONV_S1) k_conv<NV, UGK_CONV_S1, XF, PK, KS, NG><<<grid, threads, 0, s>>>(d);
  else k_conv<NV, UGK_CONV_S2, XF, PK, KS, NG><<<grid, threads, 0, s>>>(d);
}
// KS >= 4 forms with as many output groups per wave as divide the groups (<= 4)
template <int XF, int KS>
static void launch_conv_wide(const ugk_conv_desc &d, unsigned tiles_x, unsigned groups, hipStream_t s)
{
  if (groups % 4 == 0) launch_conv_nv<1, XF, 0, KS, 4>(d, dim3(tiles_x, groups / 4), s);
  else if (groups % 2 == 0) launch_conv_nv<1, XF, 0, KS, 2>(d, dim3(tiles_x, groups / 2), s);
  else launch_conv_nv<1, XF, 0, KS, 1>(d, dim3(tiles_x, groups), s);
}

// below these many blocks the coarse-grid forms (more waves per tile) run
#define CONV_WIDE_MAX 1024
#define T2_WIDE_MAX 8192
template <int XF>
static void launch_conv(const ugk_conv_desc &d, long total, hipStream_t s)
{
  const unsigned groups = (unsigned)(d.cop / 16);
  if (d.mode == UGK_CONV_T2) {
    // few groups of 16 items (the coarse grids): the chunks of the sum over
    // the waves of a block
    const long items = (vox(d.gi) + 15) / 16 * 16 * 8 * d.nb;
    if (d.cin >= 64 && items / 16 * groups < T2_WIDE_MAX)
      k_conv_t2<XF, 4><<<dim3((unsigned)(items / 16), groups), 256, 0, s>>>(d);
    else
      k_conv_t2<XF, 0><<<dim3((unsigned)((items + 63) / 64), groups), 256, 0, s>>>(d);
    return;
  }
  if (d.mode == UGK_CONV_S1 && d.cout <= 4 && XF == 1 && d.cin <= SO_MAX_CIN && d.cin % 16 == 0) {
    const unsigned nblk = (unsigned)((total + 63) / 64 < 8192 ? (total + 63) / 64 : 8192);
    if (d.cout == 3 && d.cin == 16) k_conv_small_out<1, 3><<<nblk, 256, 0, s>>>(d);   // the network's last layer
    else switch (d.cin / 16) {
      case 1: k_conv_small_out<1, 4><<<nblk, 256, 0, s>>>(d); break;
      case 2: k_conv_small_out<2, 4><<<nblk, 256, 0, s>>>(d); break;
      case 3: k_conv_small_out<3, 4><<<nblk, 256, 0, s>>>(d); break;
      default: k_conv_small_out<4, 4><<<nblk, 256, 0, s>>>(d); break;
    }
    return;
  }
  static int nv_env = -1, ks_env = -1;
  if (nv_env < 0) {
    const char *e = getenv("UFEM_CONV_NV");
    nv_env = e ? atoi(e) : 0;
    e = getenv("UFEM_CONV_KS");     /* 4, 1, or 9 for the unrolled form; unset: by size */
    ks_env = e ? atoi(e) : 0;
  }
  // unpacked inputs: the sum split over the 4 waves of a block (measured
  // fastest at every layer size of the network)
  int ks = 0;
  if (d.cin != 4) ks = (ks_env == 1 || ks_env == 9) ? (ks_env == 1 ? 1 : 0) : 4;
  const long tiles = (total + 15) / 16 * groups;
  if (ks == 4 && tiles < CONV_WIDE_MAX && (d.cin == 64 || d.cin == 32)) {
    // few tiles (the coarse levels): one wave per (chunk, plane) unit
    // (one output group per wave: on these grids more waves beat fewer, longer ones)
    if (d.cin == 64) launch_conv_nv<1, XF, 0, 12>(d, dim3((unsigned)((total + 15) / 16), groups), s);
    else launch_conv_nv<1, XF, 0, 6>(d, dim3((unsigned)((total + 15) / 16), groups), s);
    return;
  }
  if (ks == 4) {
    launch_conv_wide<XF, 4>(d, (unsigned)((total + 15) / 16), groups, s);
    return;
  }
  if (ks == 1) {
    launch_conv_nv<1, XF, 0, 1>(d, dim3((unsigned)((total + 63) / 64), groups), s);
    return;
  }
  // two voxel tiles per wave only when there are waves to spare
  const bool two = nv_env ? nv_env == 2 : total * groups >= 64L * 4 * 2048;
  const dim3 g1((unsigned)((total + 63) / 64), groups), g2((unsigned)((total + 127) / 128), groups);
  if (d.cin == 4) {
    if (two) launch_conv_nv<2, XF, 1>(d, g2, s);
    else launch_conv_nv<1, XF, 1>(d, g1, s);
  } else {
    if (two) launch_conv_nv<2, XF, 0>(d, g2, s);
    else launch_conv_nv<1, XF, 0>(d, g1, s);
  }
}

// --------------------------------------------------------------------------
// weight gradient: block (slice, tap) accumulates dW[tap] tiles over its
// voxel chunks; TPW output tiles of 16 x 16 per wave
// --------------------------------------------------------------------------
#define DW_PITCH 68      // bf16 per LDS row: 64 voxels + 4 (8-byte aligned rows)
#define DW_MAX_COP 64
#define DW_MAX_CIP 96

// 4 channels of a bf16 or f32 field as bf16 (f32 rounded to nearest even)
template <int F32>
__device__ __forceinline__ bf16x4 load4_bf(const void *p, long off)
{
  if (F32) {
    const f32x4 x = *(const f32x4 *)((const float *)p + off);
    return (bf16x4){(short)f2bf(x[0]), (short)f2bf(x[1]), (short)f2bf(x[2]), (short)f2bf(x[3])};
  }
  return *(const bf16x4 *)((const bf16 *)p + off);
}

template <int TPW, int MODE, int XF, int DF>
__global__ __launch_bounds__(256) void k_conv_dw(ugk_dw_desc d, float *part, long cps)
{
  __shared__ __attribute__((aligned(16))) bf16 sdy[DW_MAX_COP * DW_PITCH];
  __shared__ __attribute__((aligned(16))) bf16 sx[DW_MAX_CIP * DW_PITCH];
  __shared__ long src[64];
  const int t = blockIdx.y;
  const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
  const long Vo = vox(d.go), Vi = vox(d.gi);
  const long total = Vo * d.nb;
  const long nch = (total + 63) / 64;
  const long ch0 = (long)blockIdx.x * cps;
  const long ch1 = ch0 + cps < nch ? ch0 + cps : nch;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, kq = lane >> 4;
  const int nci = d.cip / 16, ntiles = (d.cop / 16) * nci;
  const int gco = d.cop / 4, gci = d.cip / 4;       // staged 4-channel groups
  const int rco = (d.cout + 3) / 4, rci = (d.cin + 3) / 4;   // groups read from memory
  const int HWo = d.go.H * d.go.W;
  f32x4 acc[TPW];
#pragma unroll
  for (int q = 0; q < TPW; q++) acc[q] = (f32x4){0.f, 0.f, 0.f, 0.f};

  for (long ch = ch0; ch < ch1; ch++) {
    const long g0 = ch * 64;
    __syncthreads();                         // previous chunk fully consumed
    if (threadIdx.x < 64) {
      const long g = g0 + threadIdx.x;
      long s = -1;
      if (g < total) {
        const long b = g / Vo;
        const int v = (int)(g - b * Vo);
        const int z = v / HWo, r = v - z * HWo, y = r / d.go.W, x = r - y * d.go.W;
        s = tap_src<MODE>(z, y, x, dz, dy, dx, d.gi);
        if (s >= 0) s += b * Vi;
      }
      src[threadIdx.x] = s;
    }
    for (int i = threadIdx.x; i < 64 * gco; i += 256) {
      const int vl = i / gco, grp = i - vl * gco;
      const long g = g0 + vl;
      bf16x4 v = (bf16x4){0, 0, 0, 0};
      if (g < total && grp < rco) v = load4_bf<DF>(d.dy, g * d.ldd + 4 * grp);
#pragma unroll
      for (int j = 0; j < 4; j++) sdy[(4 * grp + j) * DW_PITCH + vl] = (bf16)v[j];
    }
    __syncthreads();                         // src[] ready
    for (int i = threadIdx.x; i < 64 * gci; i += 256) {
      const int vl = i / gci, grp = i - vl * gci;
      const long s = src[vl];
      bf16x4 v = (bf16x4){0, 0, 0, 0};
      if (s >= 0 && grp < rci) v = load4_bf<XF>(d.x, s * d.ldx + 4 * grp);
#pragma unroll
      for (int j = 0; j < 4; j++) sx[(4 * grp + j) * DW_PITCH + vl] = (bf16)v[j];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < TPW; q++) {
      const int tile = wave + 4 * q;
      if (tile < ntiles) {                   // wave-uniform
        const int mt = tile / nci, nt = tile - mt * nci;
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
          const bf16x4 a = *(const bf16x4 *)(sdy + (mt * 16 + col) * DW_PITCH + ks * 16 + 4 * kq);
          const bf16x4 b = *(const bf16x4 *)(sx + (nt * 16 + col) * DW_PITCH + ks * 16 + 4 * kq);
          acc[q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[q], 0, 0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < TPW; q++) {
    const int tile = wave + 4 * q;
    if (tile < ntiles) {
      const int mt = tile / nci, nt = tile - mt * nci;
      float *pp = part + ((long)blockIdx.x * 27 + t) * d.cop * d.cip;
#pragma unroll
      for (int j = 0; j < 4; j++) pp[(long)(mt * 16 + 4 * kq + j) * d.cip + nt * 16 + col] = acc[q][j];
    }
  }
}

__global__ void k_dw_reduce(const float *__restrict__ part, int ks, int cop, int cip, int cout, int cin,
                            float *__restrict__ grad, int accumulate)
{
  const long P = 27L * cout * cin;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < P; i += (long)gridDim.x * blockDim.x) {
    const int t = (int)(i / ((long)cout * cin));
    const int r = (int)(i - (long)t * cout * cin);
    const int co = r / cin, ci = r - co * cin;
    float s = 0.f;
    for (int k = 0; k < ks; k++) s += part[(((long)k * 27 + t) * cop + co) * cip + ci];
    grad[i] = accumulate ? grad[i] + s : s;
  }
}

// --------------------------------------------------------------------------
// gates, loss, optimiser, training pairs
// --------------------------------------------------------------------------
__device__ __forceinline__ double block_sum(double v)
{
  __shared__ double sh[16];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += sh[w];
  return s;                                   // valid in thread 0
}

#define GRID_STRIDE(i, n) \
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (long)gridDim.x * blockDim.x)

__global__ void k_mul(long n, const float *__restrict__ a, const float *__restrict__ b, float *__restrict__ out,
                      const int *skip)
{
  if (skip && *skip) return;
  GRID_STRIDE(i, n) out[i] = a[i] * b[i];
}

// g[v][c] = exp(bp[c] - (1 + ap[c]) eb[v]) over nvox voxels (all samples)
__global__ void k_gates(long nvox, int C, const float *__restrict__ eb, const float *__restrict__ ap,
                        const float *__restrict__ bp, float *__restrict__ g)
{
  GRID_STRIDE(i, nvox * C) {
    const long v = i / C;
    const int c = (int)(i - v * C);
    g[i] = expf(bp[c] - (1.f + ap[c]) * eb[v]);
  }
}

// per-block, per-channel sums of q = g h dt and of eb q (C divides 256, so a
// thread keeps one channel over the grid-stride loop); part [blocks][2C]
__global__ __launch_bounds__(256) void k_gate_grad(long nvox, int C, const float *__restrict__ g,
                                                   const float *__restrict__ h, const float *__restrict__ dt,
                                                   const float *__restrict__ eb, double *part)
{
  __shared__ double s0[256], s1[256];
  double a = 0.0, b = 0.0;
  GRID_STRIDE(i, nvox * C) {
    const double q = (double)g[i] * h[i] * dt[i];
    a += q;
    b += q * eb[i / C];
  }
  s0[threadIdx.x] = a;
  s1[threadIdx.x] = b;
  __syncthreads();
  if ((int)threadIdx.x < C) {
    double x = 0.0, y = 0.0;
    for (int k = threadIdx.x; k < 256; k += C) {
      x += s0[k];
      y += s1[k];
    }
    part[(long)blockIdx.x * 2 * C + threadIdx.x] = x;
    part[(long)blockIdx.x * 2 * C + C + threadIdx.x] = y;
  }
}

// fixed-order sum over blocks: d/db = sum q, d/da = -sum eb q
__global__ void k_gate_reduce(int G, int C, const double *__restrict__ part, float *ga, float *gb)
{
  const int c = threadIdx.x;
  if (c >= C) return;
  double x = 0.0, y = 0.0;
  for (int k = 0; k < G; k++) {
    x += part[(long)k * 2 * C + c];
    y += part[(long)k * 2 * C + C + c];
  }
  gb[c] = (float)x;
  ga[c] = (float)(-y);
}

// energy loss of one pair: z = dinv rr + mask sqrt(dinv) y
__global__ void k_el_z(long n, const float *__restrict__ rr, const float *__restrict__ dinv,
                       const float *__restrict__ mask, const float *__restrict__ y, float *__restrict__ zz)
{
  GRID_STRIDE(i, n) {
    const long v = i / 3;
    const float d = dinv[i];
    zz[i] = d * rr[i] + mask[i] * sqrtf(d) * y[v * 4 + (i - 3 * v)];
  }
}

// res = rr - K z; loss partial sum (ee - zz) res; dy = coef mask sqrt(dinv) res
__global__ __launch_bounds__(256) void k_el_grad(long n, const float *__restrict__ rr,
                                                 const float *__restrict__ ee, const float *__restrict__ zz,
                                                 const float *__restrict__ kz, const float *__restrict__ dinv,
                                                 const float *__restrict__ mask, float coef,
                                                 float *__restrict__ dy, double *lpart)
{
  double l = 0.0;
  GRID_STRIDE(i, n) {
    const float res = rr[i] - kz[i];
    l += (double)(ee[i] - zz[i]) * (double)res;
    const long v = i / 3;
    dy[v * 4 + (i - 3 * v)] = coef * mask[i] * sqrtf(dinv[i]) * res;
  }
  const double s = block_sum(l);
  if (threadIdx.x == 0) lpart[blockIdx.x] = s;
}

__global__ void k_adam(long P, int cout, int cin, int cop, int cip, float *w, const float *__restrict__ g,
                       float *m, float *v, bf16 *wf, bf16 *wt, float lr, float b1, float b2, float eps,
                       float bc1, float bc2)
{
  GRID_STRIDE(i, P) {
    float wi = w[i];
    if (lr > 0.f) {
      const float gi = g[i];
      const float mi = b1 * m[i] + (1.f - b1) * gi;
      const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
      m[i] = mi;
      v[i] = vi;
      wi -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
      w[i] = wi;
    }
    if (wf) {
      const int t = (int)(i / ((long)cout * cin));
      const int r = (int)(i - (long)t * cout * cin);
      const int co = r / cin, ci = r - co * cin;
      const bf16 h = f2bf(wi);
      wf[((long)t * cop + co) * cip + ci] = h;
      wt[((long)t * cip + ci) * cop + co] = h;
    }
  }
}

// training pair from a PCG snapshot (x, r) and the converged u:
// partials (sum dinv r^2, sum (u - x) r) per block
__global__ __launch_bounds__(256) void k_pair_stats(long n, const double *__restrict__ x,
                                                    const double *__restrict__ r, const double *__restrict__ u,
                                                    const double *__restrict__ dinv, double *part)
{
  double a2 = 0.0, er = 0.0;
  GRID_STRIDE(i, n) {
    a2 += dinv[i] * r[i] * r[i];
    er += (u[i] - x[i]) * r[i];
  }
  const double s0 = block_sum(a2);
  const double s1 = block_sum(er);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s0;
    part[2 * blockIdx.x + 1] = s1;
  }
}

// v[vox][c] = mask sqrt(dinv) r / s, rr = r / s, ee = (u - x) / s
__global__ void k_pair_write(long n, const double *__restrict__ x, const double *__restrict__ r,
                             const double *__restrict__ u, const double *__restrict__ dinv,
                             const float *__restrict__ mask, double inv_s, float *V, float *rr, float *ee)
{
  GRID_STRIDE(i, n) {
    const long v = i / 3;
    V[v * 4 + (i - 3 * v)] = (float)(mask[i] * sqrt(dinv[i]) * r[i] * inv_s);
    rr[i] = (float)(r[i] * inv_s);
    ee[i] = (float)((u[i] - x[i]) * inv_s);
  }
}

template <int TPW, int MODE>
static void dw_launch(const ugk_dw_desc &d, dim3 grid, float *part, long cps, hipStream_t s)
{
  if (d.x_f32 && d.dy_f32) k_conv_dw<TPW, MODE, 1, 1><<<grid, 256, 0, s>>>(d, part, cps);
  else if (d.x_f32) k_conv_dw<TPW, MODE, 1, 0><<<grid, 256, 0, s>>>(d, part, cps);
  else if (d.dy_f32) k_conv_dw<TPW, MODE, 0, 1><<<grid, 256, 0, s>>>(d, part, cps);
  else k_conv_dw<TPW, MODE, 0, 0><<<grid, 256, 0, s>>>(d, part, cps);
}

__global__ __launch_bounds__(256) void k_to_split(long n4, const float *__restrict__ x, bf16 *__restrict__ y)
{
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n4; i += (long)gridDim.x * 256) {
    bf16x4 b[3];
    split3(*(const f32x4 *)(x + 4 * i), true, b);
    *(bf16x4 *)(y + 12 * i) = b[0];
    *(bf16x4 *)(y + 12 * i + 4) = b[1];
    *(bf16x4 *)(y + 12 * i + 8) = b[2];
  }
}

template <int TPW>
static void dw_mode(const ugk_dw_desc &d, dim3 grid, float *part, long cps, hipStream_t s)
{
  if (d.mode == UGK_CONV_S1) dw_launch<TPW, UGK_CONV_S1>(d, grid, part, cps, s);
  else dw_launch<TPW, UGK_CONV_S2>(d, grid, part, cps, s);
}

static unsigned grid_for(long work, int per = 256, long cap = 4096)
{
  long b = (work + per - 1) / per;
  return (unsigned)(b < 1 ? 1 : (b > cap ? cap : b));
}


// ---------------------------------------------------------------------------
// FP64 inference: the solver's application of the network
// ---------------------------------------------------------------------------
// The operator is the matrix-core path's (bf16 weights, widened exactly; f32
// gates, a fixed diagonal scaling and so part of the operator), applied with
// FP64 activations and sums.  Flexible CG with one kept direction needs the
// preconditioner to be one linear map from iteration to iteration: rounding
// noise as small as f32's (1e-7 relative) costs it its global conjugacy and
// iterations (a 24,336-DOF beam: 189 with FP64 activations, 231 to 243 with
// f32; the 1M-DOF cantilever: 284 against 331 to 427).  At 16 channels the
// network is 0.4 GFLOP, so FP64 costs little beside the launches' latency.

// FP64 copy of one layer's bf16 weights [27][cop][cip] as wu [27][co][ci]
// (the last adjoint stage, y = C0^T t0, on the vector ALUs)
__global__ void k_w64(const bf16 *__restrict__ wf, int cop, int cip, int co, int ci, double *__restrict__ wu)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 27 * co * ci) return;
  const int t = i / (co * ci), o = (i / ci) % co, c = i % ci;
  wu[i] = (double)bf2f(wf[((long)t * cop + o) * cip + c]);
}

// The 16- and 32-channel stages run on the FP64 matrix cores:
// v_mfma_f64_16x16x4f64, D (16 x 16) += A (16 x 4) B (4 x 16); lane l holds
// A[l % 16][l / 16], B[l / 16][l % 16] and D[4 v + l / 16][l % 16], v < 4.
// A wave takes 16 nodes (rows) and 16 output channels (columns); per tap the
// K dimension runs over the input channels in S = ceil(cin / 4) steps, and
// lane group g = l / 16 holds the channels g S .. g S + S - 1 (one
// contiguous gather per tap; any fixed channel-to-step map will do, the
// weights follow it).  Fragments: for output tile nt, tap t, step s, lane l
// holds W(k, j), j = 16 nt + l % 16, k = (l / 16) S + s (0 past the channels);
// the convolution's W(k, j) = W[t][j][k], its adjoint's W(k, j) = W[t][k][j].
typedef double mf64x4 __attribute__((ext_vector_type(4)));

__global__ void k_w64_frag(const bf16 *__restrict__ wf, int cop, int cip, int co, int ci, int up,
                           double *__restrict__ frag)
{
  const int nin = up ? co : ci, nout = up ? ci : co, S = (nin + 3) / 4, NT = (nout + 15) / 16;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NT * 27 * S * 64) return;
  const int l = i % 64, st = (i / 64) % S, t = (i / (64 * S)) % 27, nt = i / (64 * S * 27);
  const int j = 16 * nt + l % 16, k = (l / 16) * S + st;
  double v = 0.0;
  if (j < nout && k < nin) {
    const int o = up ? k : j, c = up ? j : k;   // the convolution's W[t][o][c]
    v = (double)bf2f(wf[((long)t * cop + o) * cip + c]);
  }
  frag[i] = v;
}

// y[c][j] = (gate[c][j] *) sum_t sum_k W[t][j][k] x[2 c + off_t][k], stride 2,
// zero padding.  A block of 4 waves and output tile blockIdx.y; with 16 or 32
// input channels the block takes 16 coarse nodes and wave w the taps
// 7 w .. 7 w + 6, all its gathers and fragment loads issued before its first
// matrix op (one latency, not 27), the four partial tiles meeting in LDS in a
// fixed order; with 3 (S = 1) each wave takes all 27 taps of its own 16 nodes.
#define N64_TPW 7
template <int S>
__global__ __launch_bounds__(256) void k_n64_down(const double *__restrict__ x, int ldx, int ci, ugk_grid3 gi,
                                                  ugk_grid3 go, const double *__restrict__ frag, int co,
                                                  const float *__restrict__ gate, double *__restrict__ y,
                                                  const int *skip)
{
  constexpr bool SPLIT = S > 1;              // the taps split over the 4 waves
  constexpr int TPW = SPLIT ? N64_TPW : 27;
  __shared__ double part[SPLIT ? 3 : 1][4][64];
  if (skip && *skip) return;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 4, c0 = g * S;
  const unsigned HWo = (unsigned)go.H * go.W, Vo = (unsigned)go.D * HWo;
  const unsigned base = SPLIT ? blockIdx.x * 16 : (blockIdx.x * 4 + w) * 16;
  const unsigned node = base + (l & 15);
  const bool nok = node < Vo;
  const unsigned nn = nok ? node : 0;
  const int cz = (int)(nn / HWo), cr = (int)(nn - (unsigned)cz * HWo), cy = cr / go.W, cx = cr - cy * go.W;
  const double *fb = frag + (long)blockIdx.y * 27 * S * 64 + l;
  double a[TPW][S], b[TPW][S];
#pragma unroll
  for (int u = 0; u < TPW; u++) {
    const int t = SPLIT ? N64_TPW * w + u : u;
    const int fz = 2 * cz + t / 9 - 1, fy = 2 * cy + (t / 3) % 3 - 1, fx = 2 * cx + t % 3 - 1;
    const bool ok = nok && t < 27 && fz >= 0 && fz < gi.D && fy >= 0 && fy < gi.H && fx >= 0 && fx < gi.W;
    const double *xs = x + ((long)(fz * gi.H + fy) * gi.W + fx) * ldx + c0;
#pragma unroll
    for (int q = 0; q < S; q++) {
      a[u][q] = (ok && c0 + q < ci) ? xs[q] : 0.0;
      b[u][q] = t < 27 ? fb[(t * S + q) * 64] : 0.0;
    }
  }
  mf64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int u = 0; u < TPW; u++)
#pragma unroll
    for (int q = 0; q < S; q++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][q], b[u][q], acc, 0, 0, 0);
  if constexpr (SPLIT) {
    if (w) {
#pragma unroll
      for (int v = 0; v < 4; v++) part[w - 1][v][l] = acc[v];
    }
    __syncthreads();
    if (w) return;
#pragma unroll
    for (int v = 0; v < 4; v++) acc[v] = ((acc[v] + part[0][v][l]) + part[1][v][l]) + part[2][v][l];
  }
  const int j = blockIdx.y * 16 + (l & 15);
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const unsigned nd = base + 4 * v + g;
    if (nd < Vo && j < co) {
      const long e = (long)nd * co + j;
      y[e] = gate ? (double)gate[e] * acc[v] : acc[v];
    }
  }
}

// the taps of class (PZ, PY, PX) of the adjoint: per axis the coarse node c
// with offset 0 (P = 0), or c with offset +1 and c + 1 with offset -1 (P = 1);
// the gathers and fragment loads of up to 32 / S taps issued before their
// matrix ops (all 8 taps of a class at 16 channels)
template <int PZ, int PY, int PX, int S>
__device__ __forceinline__ mf64x4 n64_up_class(const double *__restrict__ x, int ci, ugk_grid3 gc, int cz, int cy,
                                               int cx, bool fok, int c0, const double *__restrict__ fb)
{
  constexpr int NY = PY + 1, NX = PX + 1, NT = (PZ + 1) * NY * NX, BT = 32 / S < NT ? 32 / S : NT;
  mf64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int u0 = 0; u0 < NT; u0 += BT) {
    double a[BT][S], b[BT][S];
#pragma unroll
    for (int uu = 0; uu < BT; uu++) {
      const int u = u0 + uu;
      const int iz = u / (NY * NX), iy = (u / NX) % NY, ix = u % NX;
      const int sz = cz + iz, sy = cy + iy, sx = cx + ix;   // offset -1 reads c + 1
      const bool ok = fok && sz < gc.D && sy < gc.H && sx < gc.W;
      const int t = 9 * (PZ ? 2 - 2 * iz : 1) + 3 * (PY ? 2 - 2 * iy : 1) + (PX ? 2 - 2 * ix : 1);
      const double *xs = x + ((long)(sz * gc.H + sy) * gc.W + sx) * ci + c0;
#pragma unroll
      for (int q = 0; q < S; q++) {
        a[uu][q] = (ok && c0 + q < ci) ? xs[q] : 0.0;
        b[uu][q] = fb[(t * S + q) * 64];
      }
    }
#pragma unroll
    for (int uu = 0; uu < BT; uu++)
#pragma unroll
      for (int q = 0; q < S; q++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[uu][q], b[uu][q], acc, 0, 0, 0);
  }
  return acc;
}

// y[f][j] = gate[f][j] h[f][j] + sum over the taps of f of W[t][k][j] x[c][k]:
// the stride-2 adjoint from gc (ci channels) to gf (co channels) with the
// gated skip of the level.  Fine node f = 2 c + p of parity class p; a block
// of 8 waves takes 16 coarse nodes, wave w the fine nodes of class w (rows:
// the 16 coarse nodes), so its taps are compile-time; output tile blockIdx.y.
template <int S>
__global__ __launch_bounds__(512) void k_n64_up(const double *__restrict__ x, int ci, ugk_grid3 gc, ugk_grid3 gf,
                                                const double *__restrict__ frag, int co,
                                                const float *__restrict__ gate, const double *__restrict__ h,
                                                double *__restrict__ y, const int *skip)
{
  if (skip && *skip) return;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 4, c0 = g * S;
  const int pz = w >> 2, py = (w >> 1) & 1, px = w & 1;
  const unsigned HWc = (unsigned)gc.H * gc.W, Vc = (unsigned)gc.D * HWc;
  const unsigned c = blockIdx.x * 16 + (l & 15);
  const bool cok = c < Vc;
  const unsigned cc = cok ? c : 0;
  const int cz = (int)(cc / HWc), cr = (int)(cc - (unsigned)cz * HWc), cy = cr / gc.W, cx = cr - cy * gc.W;
  const bool fok = cok && 2 * cz + pz < gf.D && 2 * cy + py < gf.H && 2 * cx + px < gf.W;
  const double *fb = frag + (long)blockIdx.y * 27 * S * 64 + l;
  mf64x4 acc;
  switch (w) {   // wave-uniform
    case 0: acc = n64_up_class<0, 0, 0, S>(x, ci, gc, cz, cy, cx, fok, c0, fb); break;
    case 1: acc = n64_up_class<0, 0, 1, S>(x, ci, gc, cz, cy, cx, fok, c0, fb); break;
    case 2: acc = n64_up_class<0, 1, 0, S>(x, ci, gc, cz, cy, cx, fok, c0, fb); break;
    case 3: acc = n64_up_class<0, 1, 1, S>(x, ci, gc, cz, cy, cx, fok, c0, fb); break;
    case 4: acc = n64_up_class<1, 0, 0, S>(x, ci, gc, cz, cy, cx, fok, c0, fb); break;
    case 5: acc = n64_up_class<1, 0, 1, S>(x, ci, gc, cz, cy, cx, fok, c0, fb); break;
    case 6: acc = n64_up_class<1, 1, 0, S>(x, ci, gc, cz, cy, cx, fok, c0, fb); break;
    default: acc = n64_up_class<1, 1, 1, S>(x, ci, gc, cz, cy, cx, fok, c0, fb); break;
  }
  const int j = blockIdx.y * 16 + (l & 15);
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const unsigned cn = blockIdx.x * 16 + 4 * v + g;
    if (cn >= Vc || j >= co) continue;
    const int nz = (int)(cn / HWc), nr = (int)(cn - (unsigned)nz * HWc), ny = nr / gc.W, nx = nr - ny * gc.W;
    const int fz = 2 * nz + pz, fy = 2 * ny + py, fx = 2 * nx + px;
    if (fz >= gf.D || fy >= gf.H || fx >= gf.W) continue;
    const long e = (((long)fz * gf.H + fy) * gf.W + fx) * co + j;
    y[e] = gate ? acc[v] + (double)gate[e] * h[e] : acc[v];
  }
}

// y[f][o] = sum over the taps of f of W[t][k][o] x[c][k], o < 3: the last
// adjoint stage (C0^T, 16 or 32 channels to the 3 DOFs of the node grid) on
// the vector ALUs, for the network on its own (the solver folds it into its
// combine).  A block of 8 waves takes 8 coarse nodes, wave w the fine nodes
// of class w, 8 lanes per node, lane l its tap l; the sums meet by lane
// shuffles.
template <int CI>
__global__ __launch_bounds__(512) void k_n64_out(const double *__restrict__ x, ugk_grid3 gc, ugk_grid3 gf,
                                                 const double *__restrict__ wu, double *__restrict__ y)
{
  const int cls = threadIdx.x >> 6, lane = threadIdx.x & 63, ls = lane & 7;
  const int pz = cls >> 2, py = (cls >> 1) & 1, px = cls & 1;
  const int ny = py + 1, nx = px + 1, nt = (pz + 1) * ny * nx;
  const unsigned c = blockIdx.x * 8 + (lane >> 3);
  const unsigned HWc = (unsigned)gc.H * gc.W;
  const bool cok = c < (unsigned)gc.D * HWc;
  const unsigned cc = cok ? c : 0;
  const int cz = (int)(cc / HWc), cr = (int)(cc - (unsigned)cz * HWc), cy = cr / gc.W, cx = cr - cy * gc.W;
  const int fz = 2 * cz + pz, fy = 2 * cy + py, fx = 2 * cx + px;
  const bool fok = cok && fz < gf.D && fy < gf.H && fx < gf.W;
  const int iz = ls / (ny * nx), iy = (ls / nx) % ny, ix = ls % nx;
  const int sz = cz + iz, sy = cy + iy, sx = cx + ix;
  const bool ok = fok && ls < nt && sz < gc.D && sy < gc.H && sx < gc.W;
  const int t = 9 * (pz ? 2 - 2 * iz : 1) + 3 * (py ? 2 - 2 * iy : 1) + (px ? 2 - 2 * ix : 1);
  const double *xs = x + ((long)(sz * gc.H + sy) * gc.W + sx) * CI;
  double acc[3] = {0.0, 0.0, 0.0};
  if (ok) {
#pragma unroll 4
    for (int k = 0; k < CI; k++) {
      const double xv = xs[k];
#pragma unroll
      for (int o = 0; o < 3; o++) acc[o] = fma(wu[((long)t * CI + k) * 3 + o], xv, acc[o]);
    }
  }
#pragma unroll
  for (int m = 4; m >= 1; m >>= 1)
#pragma unroll
    for (int o = 0; o < 3; o++) acc[o] += __shfl_xor(acc[o], m);
  if (ls == 0 && fok) {
    double *yo = y + (((long)fz * gf.H + fy) * gf.W + fx) * 3;
#pragma unroll
    for (int o = 0; o < 3; o++) yo[o] = acc[o];
  }
}

static bool n64_grid_ok(ugk_grid3 gi, ugk_grid3 go)
{
  return go.D == (gi.D + 1) / 2 && go.H == (gi.H + 1) / 2 && go.W == (gi.W + 1) / 2 && vox(gi) < (1L << 31);
}

extern "C" {

int ugk_conv(const ugk_conv_desc *d, void *s)
{
  const long total = vox(d->go) * d->nb;
  if (total <= 0) return 0;
  if (d->cin % 4 || d->cip % 16 || d->cop % 16 || d->cin > d->cip || d->cout > d->cop || d->ldx % 4 ||
      d->ldy % 4 || d->mode < 0 || d->mode > 2 || (d->ga && !d->gb))
    return ugk_hip_fail("ugk_conv: channel layout", hipErrorInvalidValue);
  if (d->x_f32 < 0 || d->x_f32 > 2 || d->y_f32 < 0 || d->y_f32 > 1)
    return ugk_hip_fail("ugk_conv: activation format", hipErrorInvalidValue);
  if (total >= (1L << 31) || (vox(d->gi) + 15) / 16 * 16 * 8 * d->nb >= (1L << 31))   // 32-bit voxel indices
    return ugk_hip_fail("ugk_conv: grid too large", hipErrorInvalidValue);
  if (d->x_f32 == UGK_ACT_SPLIT) launch_conv<2>(*d, total, S(s));
  else if (d->x_f32 == UGK_ACT_F32) launch_conv<1>(*d, total, S(s));
  else launch_conv<0>(*d, total, S(s));
  CK_LAUNCH("k_conv");
  return 0;
}

int ugk_conv_dw_slices(const ugk_dw_desc *d)
{
  const long nch = (vox(d->go) * d->nb + 63) / 64;
  long ks = nch / 8;
  return (int)(ks < 1 ? 1 : (ks > 96 ? 96 : ks));
}

int ugk_conv_dw(const ugk_dw_desc *d, float *part, int ks, float *grad, int accumulate, void *s)
{
  const long nch = (vox(d->go) * d->nb + 63) / 64;
  if (nch <= 0) return 0;
  if (d->cop > DW_MAX_COP || d->cip > DW_MAX_CIP || d->cop % 16 || d->cip % 16 || d->ldx % 4 ||
      d->ldd % 4 || (d->mode != UGK_CONV_S1 && d->mode != UGK_CONV_S2) || ks < 1 ||
      d->x_f32 < 0 || d->x_f32 > 1 || d->dy_f32 < 0 || d->dy_f32 > 1)
    return ugk_hip_fail("ugk_conv_dw: layout", hipErrorInvalidValue);
  const long cps = (nch + ks - 1) / ks;
  const int tiles = (d->cop / 16) * (d->cip / 16);
  dim3 grid((unsigned)((nch + cps - 1) / cps), 27);
  const int ksr = (int)grid.x;
  if (tiles <= 4) dw_mode<1>(*d, grid, part, cps, S(s));
  else if (tiles <= 8) dw_mode<2>(*d, grid, part, cps, S(s));
  else if (tiles <= 16) dw_mode<4>(*d, grid, part, cps, S(s));
  else if (tiles <= 24) dw_mode<6>(*d, grid, part, cps, S(s));
  else return ugk_hip_fail("ugk_conv_dw: too many tiles", hipErrorInvalidValue);
  CK_LAUNCH("k_conv_dw");
  k_dw_reduce<<<grid_for(27L * d->cout * d->cin), 256, 0, S(s)>>>(part, ksr, d->cop, d->cip, d->cout, d->cin,
                                                                   grad, accumulate);
  CK_LAUNCH("k_dw_reduce");
  return 0;
}

int ugk_to_split(long n4, const float *x, unsigned short *y, void *s)
{
  if (n4 <= 0) return 0;
  k_to_split<<<grid_for(n4), 256, 0, S(s)>>>(n4, x, y);
  CK_LAUNCH("k_to_split");
  return 0;
}

int ugk_mul(long n, const float *a, const float *b, float *out, const int *skip, void *s)
{
  k_mul<<<grid_for(n), 256, 0, S(s)>>>(n, a, b, out, skip);
  CK_LAUNCH("k_mul");
  return 0;
}

int ugk_gates(long nvox, int C, const float *eb, const float *ap, const float *bp, float *g, void *s)
{
  k_gates<<<grid_for(nvox * C), 256, 0, S(s)>>>(nvox, C, eb, ap, bp, g);
  CK_LAUNCH("k_gates");
  return 0;
}

int ugk_gate_grad(long nvox, int C, const float *g, const float *h, const float *dt, const float *eb,
                  double *part, float *grad_a, float *grad_b, void *s)
{
  if (C < 1 || C > 256 || 256 % C) return ugk_hip_fail("ugk_gate_grad: channels", hipErrorInvalidValue);
  const unsigned G = grid_for(nvox * C, 256, 512);
  k_gate_grad<<<G, 256, 0, S(s)>>>(nvox, C, g, h, dt, eb, part);
  CK_LAUNCH("k_gate_grad");
  k_gate_reduce<<<1, 256, 0, S(s)>>>((int)G, C, part, grad_a, grad_b);
  CK_LAUNCH("k_gate_reduce");
  return 0;
}

int ugk_el_z(long n, const float *rr, const float *dinv, const float *mask, const float *y, float *zz, void *s)
{
  k_el_z<<<grid_for(n), 256, 0, S(s)>>>(n, rr, dinv, mask, y, zz);
  CK_LAUNCH("k_el_z");
  return 0;
}

int ugk_el_grad(long n, const float *rr, const float *ee, const float *zz, const float *kz, const float *dinv,
                const float *mask, float coef, float *dy, double *lpart, void *s)
{
  const unsigned G = grid_for(n, 256, 1024);
  k_el_grad<<<G, 256, 0, S(s)>>>(n, rr, ee, zz, kz, dinv, mask, coef, dy, lpart);
  CK_LAUNCH("k_el_grad");
  return (int)G;
}

int ugk_adam(long P, int cout, int cin, int cop, int cip, float *w, const float *g, float *m, float *v,
             unsigned short *wf, unsigned short *wt, float lr, float b1, float b2, float eps, long step,
             void *s)
{
  const float bc1 = step > 0 ? 1.f - powf(b1, (float)step) : 1.f;
  const float bc2 = step > 0 ? 1.f - powf(b2, (float)step) : 1.f;
  k_adam<<<grid_for(P), 256, 0, S(s)>>>(P, cout, cin, cop, cip, w, g, m, v, wf, wt, lr, b1, b2, eps, bc1, bc2);
  CK_LAUNCH("k_adam");
  return 0;
}

int ugk_pair_stats(long n, const double *x, const double *r, const double *u, const double *dinv, double *part,
                   void *s)
{
  const unsigned G = grid_for(n, 256, 512);
  k_pair_stats<<<G, 256, 0, S(s)>>>(n, x, r, u, dinv, part);
  CK_LAUNCH("k_pair_stats");
  return (int)G;
}

int ugk_pair_write(long n, const double *x, const double *r, const double *u, const double *dinv,
                   const float *mask, double sc, float *V, float *rr, float *ee, void *s)
{
  k_pair_write<<<grid_for(n), 256, 0, S(s)>>>(n, x, r, u, dinv, mask, sc > 0 ? 1.0 / sc : 0.0, V, rr, ee);
  CK_LAUNCH("k_pair_write");
  return 0;
}

int ugk_w64(const unsigned short *wf, int cop, int cip, int co, int ci, double *wu, void *s)
{
  const int n = 27 * co * ci;
  k_w64<<<(n + 255) / 256, 256, 0, S(s)>>>(wf, cop, cip, co, ci, wu);
  CK_LAUNCH("k_w64");
  return 0;
}

long ugk_n64_frag_len(int nin, int nout) { return (long)((nout + 15) / 16) * 27 * ((nin + 3) / 4) * 64; }

int ugk_w64_frag(const unsigned short *wf, int cop, int cip, int co, int ci, int up, double *frag, void *s)
{
  const long n = up ? ugk_n64_frag_len(co, ci) : ugk_n64_frag_len(ci, co);
  k_w64_frag<<<(unsigned)((n + 255) / 256), 256, 0, S(s)>>>(wf, cop, cip, co, ci, up, frag);
  CK_LAUNCH("k_w64_frag");
  return 0;
}

int ugk_n64_down(const double *x, int ldx, int ci, ugk_grid3 gi, ugk_grid3 go, const double *frag, int co,
                 const float *gate, double *y, const int *skip, void *s)
{
  if (!n64_grid_ok(gi, go) || ldx < ci || (co != 16 && co != 32) || (ci != 3 && ci != 16 && ci != 32))
    return ugk_hip_fail("ugk_n64_down: channels or grids", hipErrorInvalidValue);
  const dim3 G((unsigned)((vox(go) + (ci == 3 ? 63 : 15)) / (ci == 3 ? 64 : 16)), (unsigned)(co / 16));
  if (ci == 3) k_n64_down<1><<<G, 256, 0, S(s)>>>(x, ldx, ci, gi, go, frag, co, gate, y, skip);
  else if (ci == 16) k_n64_down<4><<<G, 256, 0, S(s)>>>(x, ldx, ci, gi, go, frag, co, gate, y, skip);
  else k_n64_down<8><<<G, 256, 0, S(s)>>>(x, ldx, ci, gi, go, frag, co, gate, y, skip);
  CK_LAUNCH("k_n64_down");
  return 0;
}

int ugk_n64_up(const double *x, int ci, ugk_grid3 gc, ugk_grid3 gf, const double *frag, int co, const float *gate,
               const double *h, double *y, const int *skip, void *s)
{
  if (!n64_grid_ok(gf, gc) || (co != 16 && co != 32) || (ci != 16 && ci != 32))
    return ugk_hip_fail("ugk_n64_up: channels or grids", hipErrorInvalidValue);
  const dim3 G((unsigned)((vox(gc) + 15) / 16), (unsigned)(co / 16));
  if (ci == 16) k_n64_up<4><<<G, 512, 0, S(s)>>>(x, ci, gc, gf, frag, co, gate, h, y, skip);
  else k_n64_up<8><<<G, 512, 0, S(s)>>>(x, ci, gc, gf, frag, co, gate, h, y, skip);
  CK_LAUNCH("k_n64_up");
  return 0;
}

int ugk_n64_out(const double *x, int ci, ugk_grid3 gc, ugk_grid3 gf, const double *wu, double *y, void *s)
{
  if (!n64_grid_ok(gf, gc) || (ci != 16 && ci != 32))
    return ugk_hip_fail("ugk_n64_out: channels or grids", hipErrorInvalidValue);
  const unsigned G = (unsigned)((vox(gc) + 7) / 8);
  if (ci == 16) k_n64_out<16><<<G, 512, 0, S(s)>>>(x, gc, gf, wu, y);
  else k_n64_out<32><<<G, 512, 0, S(s)>>>(x, gc, gf, wu, y);
  CK_LAUNCH("k_n64_out");
  return 0;
}


}  // extern "C"
