// Waveguide mode solver on the device: the subspace iteration of pjz's
// mode() (src/pjz/_mode.py:99-142) over the operator of a 2-D cross-section
// (_operator, src/pjz/_mode.py:22-50), run to convergence inside one launch,
// so the host never synchronises inside the loop (a torch version of the same
// loop is bound by ~30 launches and one host read per iteration).
//
// Per iteration, four passes over the vectors:
//   1. x = y R1^-1, summing the Gram matrix of x as the rows are written;
//   2. x = x R2^-1 (1 and 2 are CholeskyQR2: Gram matrices in double,
//      Cholesky and the triangular inverse on one lane);
//   3. y = A x, first stencil pass: the scalar fields b = (-(d-y u0) +
//      d-x u1) / eps_z and c = d+x u0 + d+y u1;
//   4. second stencil pass, the vector update, summing w_k = <x_k, y_k>,
//      |y_k|^2, |x_k|^2 (so err_k = |y_k - w_k x_k|) and the Gram matrix of
//      y, which gives the next iteration's R1.
// Vectors are (2, XX, YY, MM) with the MM modes innermost, as pjz lays them
// out; the periodic differences wrap in x and y (jnp.roll in pjz).  Every
// pass is a stride loop over rows or cells, so any cross-section size works;
// fields live in global memory and the sums go through LDS.
//
// Two kernels run that iteration:
//  - k_mode_subspace: one workgroup per frequency; the frequency stops once
//    its residuals are all <= tol.  Small sections (a few thousand cells),
//    where barriers and L2 latency dominate.
//  - k_mode_grid: `gpf` workgroups per frequency in one cooperative launch,
//    the passes split across them and separated by grid barriers, partial
//    sums exchanged through a workspace; every workgroup reduces the partials
//    itself in a fixed order, so all hold the same R^-1 and the same stop
//    decision.  All frequencies iterate together until the largest residual
//    of all is <= tol, which is pjz's own test (it takes the max over
//    frequencies and modes).  Large sections, where one CU per frequency
//    bounds the first kernel.
#include <hip/hip_cooperative_groups.h>

#include "fdtd_common.h"

namespace pjz {

namespace cg = cooperative_groups;

constexpr int kModeThreads = 512;
constexpr int kModeWaves = kModeThreads / 64;
constexpr int kModeMaxShare = 32;  // grid kernel: most workgroups per frequency

struct ModeArgs {
  int XX, YY, max_iters;
  float tol;
  const float* eps;    // (3, XX, YY): eps_x, eps_y, eps_z (propagation along z)
  const float* omega;  // (ww)
  const float* shift;  // (ww)
  float* x;            // (ww, 2, XX, YY, MM): start vectors in, final basis out
  float* y;            // workspace, like x: A x
  float* bc;           // workspace, like x: the b and c scalar fields
  float* w;            // (ww, MM)
  float* err;          // (ww, MM)
  int* iters;          // (ww)
  double* partial;     // grid kernel: (ww, gpf, NQ + NPAIR) partial sums
  int gpf;             // grid kernel: workgroups per frequency
};

template <int MM>
struct ModeLds {
  static constexpr int NPAIR = MM * (MM + 1) / 2;  // Gram entries (upper triangle)
  static constexpr int NQ = 3 * MM;                // <x,y>, <y,y>, <x,x> per mode
  static constexpr int NALL = NQ + NPAIR;
  double red[kModeWaves][NALL];
  double sum[NALL];    // [NQ Rayleigh sums | NPAIR Gram entries]
  double R[MM * MM];   // Cholesky factor, upper (G = R^T R)
  double Ri[MM * MM];  // its inverse, upper
  int go;
};

// Sums acc[0..P) over the workgroup into out[0..P) (LDS): wavefront
// reductions, one LDS row per wave, one barrier, the cross-wave sums, one
// barrier.  Fixed order, so results do not depend on timing.
template <int MM, int P>
__device__ __forceinline__ void block_sum(const double (&acc)[P], ModeLds<MM>& L, double* out) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int p = 0; p < P; ++p) {
    double v = acc[p];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) L.red[wv][p] = v;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < P; q += kModeThreads) {
    double t = 0.0;
    for (int w = 0; w < kModeWaves; ++w) t += L.red[w][q];
    out[q] = t;
  }
  __syncthreads();
}

// Index of Gram entry (a, b), a <= b, in the row-by-row upper triangle.
template <int MM>
__device__ constexpr int tri(int a, int b) { return a * MM - a * (a - 1) / 2 + (b - a); }

// acc[off + tri(a, b)] += v[a] v[b] for a <= b (the loops unroll completely,
// so every register index is a constant).
template <int MM, int P>
__device__ __forceinline__ void add_gram(double (&acc)[P], int off, const float (&v)[MM]) {
#pragma unroll
  for (int a = 0; a < MM; ++a) {
#pragma unroll
    for (int b = a; b < MM; ++b) acc[off + tri<MM>(a, b)] += (double)v[a] * (double)v[b];
  }
}

// On one lane: R with G = R^T R for the Gram matrix in g (upper triangle),
// and Ri = R^-1, into LDS.  The caller synchronises.
template <int MM>
__device__ __forceinline__ void chol_inv(const double* g, ModeLds<MM>& L) {
  double* R = L.R;
  double* Ri = L.Ri;
  for (int i = 0; i < MM * MM; ++i) R[i] = Ri[i] = 0.0;
  for (int j = 0; j < MM; ++j) {
    double d = g[tri<MM>(j, j)];
    for (int k = 0; k < j; ++k) d -= R[k * MM + j] * R[k * MM + j];
    // A start vector inside the span of the others (degenerate input only)
    // keeps a tiny pivot instead of a NaN.
    R[j * MM + j] = sqrt(d > 1e-300 ? d : 1e-300);
    for (int i = j + 1; i < MM; ++i) {
      double t = g[tri<MM>(j, i)];
      for (int k = 0; k < j; ++k) t -= R[k * MM + j] * R[k * MM + i];
      R[j * MM + i] = t / R[j * MM + j];
    }
  }
  for (int j = 0; j < MM; ++j) {
    Ri[j * MM + j] = 1.0 / R[j * MM + j];
    for (int i = j - 1; i >= 0; --i) {
      double t = 0.0;
      for (int k = i + 1; k <= j; ++k) t += R[i * MM + k] * Ri[k * MM + j];
      Ri[i * MM + j] = -t / R[i * MM + i];
    }
  }
}

// One row times Ri (upper triangular).
template <int MM>
__device__ __forceinline__ void row_ri(float (&v)[MM], const ModeLds<MM>& L) {
  float o[MM];
#pragma unroll
  for (int k = 0; k < MM; ++k) {
    double t = 0.0;
#pragma unroll
    for (int j = 0; j <= k; ++j) t += (double)v[j] * L.Ri[j * MM + k];
    o[k] = (float)t;
  }
#pragma unroll
  for (int k = 0; k < MM; ++k) v[k] = o[k];
}

// A thread's share of a pass: items t0, t0 + stride, ... (one workgroup:
// threadIdx.x and the block size; the grid kernel: the frequency's threads).
struct Share {
  int t0, stride;
};

// acc[off..] += Gram matrix of the rows of v in this share.
template <int MM, int P>
__device__ __forceinline__ void gram_rows(const float* __restrict__ v, int n, Share sh,
                                          double (&acc)[P], int off) {
  for (int r = sh.t0; r < n; r += sh.stride) {
    float row[MM];
#pragma unroll
    for (int k = 0; k < MM; ++k) row[k] = v[r * MM + k];
    add_gram<MM>(acc, off, row);
  }
}

// dst = src Ri for the rows of this share (dst may alias src); with GRAM the
// Gram matrix of the new rows goes into acc.
template <int MM, bool GRAM, int P>
__device__ __forceinline__ void ri_rows(const float* src, float* dst, int n, Share sh,
                                        const ModeLds<MM>& L, double (&acc)[P]) {
  for (int r = sh.t0; r < n; r += sh.stride) {
    float v[MM];
#pragma unroll
    for (int k = 0; k < MM; ++k) v[k] = src[r * MM + k];
    row_ri<MM>(v, L);
#pragma unroll
    for (int k = 0; k < MM; ++k) dst[r * MM + k] = v[k];
    if constexpr (GRAM) add_gram<MM>(acc, 0, v);
  }
}

// First stencil pass of y = A x (_operator, src/pjz/_mode.py:22-50): the
// scalar fields b = (-(d-y u0) + d-x u1) / eps_z (backward differences) and
// c = d+x u0 + d+y u1 (forward differences), periodic in x and y.
template <int MM>
__device__ __forceinline__ void op_pass1(const float* __restrict__ x, float* __restrict__ bc,
                                         const float* __restrict__ eps, int XX, int YY,
                                         Share sh) {
  const int P = XX * YY, cells = P * MM;
  const float* ez = eps + 2 * P;
  for (int e = sh.t0; e < cells; e += sh.stride) {
    const int p = e / MM, k = e - p * MM;
    const int i = p / YY, j = p - i * YY;
    const int im = i == 0 ? XX - 1 : i - 1, ip = i == XX - 1 ? 0 : i + 1;
    const int jm = j == 0 ? YY - 1 : j - 1, jp = j == YY - 1 ? 0 : j + 1;
    const float* x0 = x;
    const float* x1 = x + cells;
    const float u0 = x0[e], u1 = x1[e];
    bc[e] = (-(u0 - x0[(i * YY + jm) * MM + k]) + (u1 - x1[(im * YY + j) * MM + k])) / ez[p];
    bc[cells + e] = (x0[(ip * YY + j) * MM + k] - u0) + (x1[(i * YY + jp) * MM + k] - u1);
  }
}

// Whether the second stencil pass also sums the Gram matrix of y: its
// 3 MM + MM (MM + 1) / 2 double accumulators fit in registers up to 6 modes;
// above, a separate pass sums the Gram matrix.
template <int MM>
constexpr bool kFuseGram = MM <= 6;

// Second stencil pass, one cell (both components, all modes) per item, with
// the same grouping of the float operations as _operator: y = A x, and into
// acc the Rayleigh sums (x.y, y.y, x.x per mode) and, when P has room for it,
// the Gram matrix of y.
template <int MM, int P>
__device__ __forceinline__ void op_pass2(const float* __restrict__ x, float* __restrict__ y,
                                         const float* __restrict__ bc,
                                         const float* __restrict__ eps, int XX, int YY,
                                         float om2, float s, Share sh, double (&acc)[P]) {
  const int PC = XX * YY, cells = PC * MM;
  const float* ex = eps;
  const float* ey = eps + PC;
  const float* bb = bc;
  const float* cc = bc + cells;
  for (int p = sh.t0; p < PC; p += sh.stride) {
    const int i = p / YY, j = p - i * YY;
    const int im = i == 0 ? XX - 1 : i - 1, ip = i == XX - 1 ? 0 : i + 1;
    const int jm = j == 0 ? YY - 1 : j - 1, jp = j == YY - 1 ? 0 : j + 1;
    const float exv = ex[p], eyv = ey[p];
    float x0[MM], x1[MM], y0[MM], y1[MM];
#pragma unroll
    for (int k = 0; k < MM; ++k) {
      const int e = p * MM + k;
      const float b = bb[e], c = cc[e];
      x0[k] = x[e];
      x1[k] = x[cells + e];
      const float a0 = (om2 * eyv - s) * x0[k], a1 = (om2 * exv - s) * x1[k];
      const float b0 = -(bb[(i * YY + jp) * MM + k] - b) * eyv;
      const float b1 = (bb[(ip * YY + j) * MM + k] - b) * exv;
      y0[k] = (a0 + b0) + (c - cc[(im * YY + j) * MM + k]);
      y1[k] = (a1 + b1) + (c - cc[(i * YY + jm) * MM + k]);
      y[e] = y0[k];
      y[cells + e] = y1[k];
      acc[k] += (double)x0[k] * y0[k] + (double)x1[k] * y1[k];
      acc[MM + k] += (double)y0[k] * y0[k] + (double)y1[k] * y1[k];
      acc[2 * MM + k] += (double)x0[k] * x0[k] + (double)x1[k] * x1[k];
    }
    if constexpr (P > 3 * MM) {
      add_gram<MM>(acc, 3 * MM, y0);
      add_gram<MM>(acc, 3 * MM, y1);
    }
  }
}

// Residuals of one frequency from its Rayleigh sums s[0..3 MM); writes w and
// err when `out`; returns max err (NaN if any residual is NaN).
template <int MM>
__device__ __forceinline__ double residuals(const double* s, float* w, float* err, bool out) {
  double emax = 0.0;
  bool nan = false;
  for (int k = 0; k < MM; ++k) {
    const double xy = s[k], yy = s[MM + k], xx = s[2 * MM + k];
    const double e2 = yy - 2.0 * xy * xy + xy * xy * xx;
    const bool bad = !(e2 == e2);
    const double e = sqrt(e2 > 0.0 ? e2 : 0.0);
    nan |= bad;
    emax = e > emax ? e : emax;
    if (out) {
      w[k] = (float)xy;
      err[k] = bad ? __builtin_nanf("") : (float)e;
    }
  }
  return nan ? __builtin_nan("") : emax;
}

// --------------------------------------------------------------- one workgroup

// The stencil's second pass with all its sums into L.sum (Rayleigh sums, then
// the Gram matrix of y), in one pass or two.
template <int MM>
__device__ __forceinline__ void op_pass2_sums(const float* __restrict__ x, float* __restrict__ y,
                                              const float* __restrict__ bc,
                                              const float* __restrict__ eps, int XX, int YY,
                                              float om2, float s, int n, ModeLds<MM>& L) {
  constexpr int NQ = ModeLds<MM>::NQ, NPAIR = ModeLds<MM>::NPAIR;
  const Share sh{(int)threadIdx.x, kModeThreads};
  if constexpr (kFuseGram<MM>) {
    double acc[NQ + NPAIR];
#pragma unroll
    for (int q = 0; q < NQ + NPAIR; ++q) acc[q] = 0.0;
    op_pass2<MM>(x, y, bc, eps, XX, YY, om2, s, sh, acc);
    block_sum<MM>(acc, L, L.sum);
  } else {
    {
      double acc[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
      op_pass2<MM>(x, y, bc, eps, XX, YY, om2, s, sh, acc);
      block_sum<MM>(acc, L, L.sum);
    }
    double acc[NPAIR];
#pragma unroll
    for (int q = 0; q < NPAIR; ++q) acc[q] = 0.0;
    gram_rows<MM>(y, n, sh, acc, 0);
    block_sum<MM>(acc, L, L.sum + NQ);
  }
}

template <int MM>
__global__ __launch_bounds__(kModeThreads) void k_mode_subspace(ModeArgs a) {
  __shared__ ModeLds<MM> L;
  constexpr int NQ = ModeLds<MM>::NQ, NPAIR = ModeLds<MM>::NPAIR;
  const int f = blockIdx.x;
  const int n = 2 * a.XX * a.YY;  // rows of the (n, MM) matrices
  const long long off = (long long)f * n * MM;
  float* x = a.x + off;
  float* y = a.y + off;
  float* bc = a.bc + off;
  const float om = a.omega[f];
  const float om2 = om * om;
  const float s = a.shift[f];
  const Share sh{(int)threadIdx.x, kModeThreads};
  // y = A x0 and its Gram matrix; then R^-1 for the first QR.
  op_pass1<MM>(x, bc, a.eps, a.XX, a.YY, sh);
  __syncthreads();
  op_pass2_sums<MM>(x, y, bc, a.eps, a.XX, a.YY, om2, s, n, L);
  if (threadIdx.x == 0) chol_inv<MM>(L.sum + NQ, L);
  __syncthreads();
  int it = 0;
  bool go = true;  // max err > tol; err starts at +inf
  while (it < a.max_iters && go) {
    // CholeskyQR2 of y: x = y R1^-1 (with the Gram matrix of that x), then
    // x = x R2^-1.
    {
      double acc[NPAIR];
#pragma unroll
      for (int q = 0; q < NPAIR; ++q) acc[q] = 0.0;
      ri_rows<MM, (MM <= 7)>(y, x, n, sh, L, acc);
      if constexpr (MM > 7) {
        // Above 7 modes the Gram sum takes a pass of its own (registers);
        // each thread reads back only the rows it wrote.
        gram_rows<MM>(x, n, sh, acc, 0);
      }
      block_sum<MM>(acc, L, L.sum + NQ);
      if (threadIdx.x == 0) chol_inv<MM>(L.sum + NQ, L);
      __syncthreads();
      double none[1];
      ri_rows<MM, false>(x, x, n, sh, L, none);
      __syncthreads();
    }
    // y = A x with the Rayleigh sums and the Gram matrix of the new y.
    op_pass1<MM>(x, bc, a.eps, a.XX, a.YY, sh);
    __syncthreads();
    op_pass2_sums<MM>(x, y, bc, a.eps, a.XX, a.YY, om2, s, n, L);
    if (threadIdx.x == 0) {
      const double emax = residuals<MM>(L.sum, a.w + f * MM, a.err + f * MM, true);
      // pjz continues while max(err) > tol, which a NaN ends.
      L.go = emax > (double)a.tol;
      // R^-1 of the new y for the next iteration's first QR step.
      chol_inv<MM>(L.sum + NQ, L);
    }
    __syncthreads();
    go = L.go;
    ++it;
  }
  if (threadIdx.x == 0) {
    a.iters[f] = it;
    if (it == 0) {
      for (int k = 0; k < MM; ++k) {
        a.w[f * MM + k] = 0.0f;
        a.err[f * MM + k] = __builtin_inff();
      }
    }
  }
}

// ------------------------------------------------------------------ the grid

// Writes P of this workgroup's sums (L.sum[off..off+P)) into its partial row.
__device__ __forceinline__ void put_partial(double* partial, long long row, int stride,
                                            const double* sums, int off, int P) {
  for (int q = threadIdx.x; q < P; q += kModeThreads)
    partial[row * stride + off + q] = sums[off + q];
}

// out[off + q] = sum over the frequency's gpf partial rows of entry off + q,
// in workgroup order, so every workgroup gets the same bits.
__device__ __forceinline__ void reduce_partials(const double* partial, int f, int gpf,
                                                int stride, int off, int P, double* out) {
  for (int q = threadIdx.x; q < P; q += kModeThreads) {
    double t = 0.0;
    for (int j = 0; j < gpf; ++j) t += partial[((long long)f * gpf + j) * stride + off + q];
    out[off + q] = t;
  }
}

// The grid kernel's second stencil pass with its sums: partial rows of all
// NQ + NPAIR sums, grid barrier, then every workgroup reduces its frequency's
// into L.sum.  On entry the previous partials are no longer read.
template <int MM>
__device__ __forceinline__ void grid_pass2(cg::grid_group& grid, const ModeArgs& a,
                                           const float* x, float* y, const float* bc, float om2,
                                           float s, int n, int f, long long prow, Share sh,
                                           ModeLds<MM>& L) {
  constexpr int NQ = ModeLds<MM>::NQ, NPAIR = ModeLds<MM>::NPAIR, NALL = ModeLds<MM>::NALL;
  if constexpr (kFuseGram<MM>) {
    double acc[NALL];
#pragma unroll
    for (int q = 0; q < NALL; ++q) acc[q] = 0.0;
    op_pass2<MM>(x, y, bc, a.eps, a.XX, a.YY, om2, s, sh, acc);
    block_sum<MM>(acc, L, L.sum);
    put_partial(a.partial, prow, NALL, L.sum, 0, NALL);
  } else {
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
    op_pass2<MM>(x, y, bc, a.eps, a.XX, a.YY, om2, s, sh, acc);
    block_sum<MM>(acc, L, L.sum);
    put_partial(a.partial, prow, NALL, L.sum, 0, NQ);
    // Rows (Gram sum) and cells (stencil) are shared out differently, so
    // all of y must be written before its Gram sum.
    grid.sync();
    double accg[NPAIR];
#pragma unroll
    for (int q = 0; q < NPAIR; ++q) accg[q] = 0.0;
    gram_rows<MM>(y, n, sh, accg, 0);
    block_sum<MM>(accg, L, L.sum + NQ);
    put_partial(a.partial, prow, NALL, L.sum, NQ, NPAIR);
  }
  grid.sync();
  reduce_partials(a.partial, f, a.gpf, NALL, 0, NALL, L.sum);
  __syncthreads();
}

template <int MM>
__global__ __launch_bounds__(kModeThreads) void k_mode_grid(ModeArgs a) {
  __shared__ ModeLds<MM> L;
  constexpr int NQ = ModeLds<MM>::NQ, NPAIR = ModeLds<MM>::NPAIR, NALL = ModeLds<MM>::NALL;
  cg::grid_group grid = cg::this_grid();
  const int gpf = a.gpf;
  const int ww = gridDim.x / gpf;
  const int f = blockIdx.x / gpf, part = blockIdx.x - f * gpf;
  const long long prow = (long long)f * gpf + part;  // this workgroup's partial row
  const int n = 2 * a.XX * a.YY;
  const long long off = (long long)f * n * MM;
  float* x = a.x + off;
  float* y = a.y + off;
  float* bc = a.bc + off;
  const float om = a.omega[f];
  const float om2 = om * om;
  const float s = a.shift[f];
  const Share sh{part * kModeThreads + (int)threadIdx.x, gpf * kModeThreads};

  op_pass1<MM>(x, bc, a.eps, a.XX, a.YY, sh);
  grid.sync();
  grid_pass2<MM>(grid, a, x, y, bc, om2, s, n, f, prow, sh, L);
  if (threadIdx.x == 0) chol_inv<MM>(L.sum + NQ, L);
  __syncthreads();
  int it = 0;
  bool go = true;  // max err > tol over all frequencies; err starts at +inf
  while (it < a.max_iters && go) {
    // x = y R1^-1 with the partial Gram sums of x.  The barrier before the
    // partials are written makes sure every workgroup has finished reading
    // the previous ones (its reduce and the stop test).
    {
      double acc[NPAIR];
#pragma unroll
      for (int q = 0; q < NPAIR; ++q) acc[q] = 0.0;
      ri_rows<MM, (MM <= 7)>(y, x, n, sh, L, acc);
      if constexpr (MM > 7) gram_rows<MM>(x, n, sh, acc, 0);  // own rows only
      block_sum<MM>(acc, L, L.sum + NQ);
      grid.sync();
      put_partial(a.partial, prow, NALL, L.sum, NQ, NPAIR);
      grid.sync();
      reduce_partials(a.partial, f, gpf, NALL, NQ, NPAIR, L.sum);
      __syncthreads();
      if (threadIdx.x == 0) chol_inv<MM>(L.sum + NQ, L);
      __syncthreads();
      double none[1];
      ri_rows<MM, false>(x, x, n, sh, L, none);
    }
    grid.sync();  // x complete before the stencil reads neighbours
    op_pass1<MM>(x, bc, a.eps, a.XX, a.YY, sh);
    grid.sync();
    grid_pass2<MM>(grid, a, x, y, bc, om2, s, n, f, prow, sh, L);
    if (threadIdx.x == 0) {
      // The stop test over all frequencies (their Rayleigh sums from the
      // partials, reduced in the same order everywhere); the own frequency's
      // w and err written by its first workgroup.
      double emax = 0.0;
      bool nan = false;
      for (int g = 0; g < ww; ++g) {
        double sums[NQ];
        for (int q = 0; q < NQ; ++q) {
          double t = 0.0;
          for (int j = 0; j < gpf; ++j) t += a.partial[((long long)g * gpf + j) * NALL + q];
          sums[q] = t;
        }
        const double e = residuals<MM>(sums, a.w + g * MM, a.err + g * MM, g == f && part == 0);
        nan |= !(e == e);
        emax = e > emax ? e : emax;
      }
      L.go = !nan && emax > (double)a.tol;
      chol_inv<MM>(L.sum + NQ, L);
    }
    __syncthreads();
    go = L.go;
    ++it;
  }
  if (threadIdx.x == 0 && part == 0) {
    a.iters[f] = it;
    if (it == 0) {
      for (int k = 0; k < MM; ++k) {
        a.w[f * MM + k] = 0.0f;
        a.err[f * MM + k] = __builtin_inff();
      }
    }
  }
}

// ------------------------------------------------------------------ host side

// Workgroups per frequency for the grid kernel, or 0 to use one workgroup:
// about one workgroup per 8192 cells x modes, at most kModeMaxShare (grid
// barriers cost more as the grid grows), and only when the whole grid fits on
// the device at once (a cooperative launch needs every workgroup resident).
template <int MM>
static int grid_share(int ww, int XX, int YY) {
  const long long work = (long long)XX * YY * MM;
  if (work < 16384) return 0;
  int gpf = (int)std::min<long long>(kModeMaxShare, (work + 8191) / 8192);
  int dev = 0, cus = 0, coop = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess ||
      !coop ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu, reinterpret_cast<const void*>(&k_mode_grid<MM>), kModeThreads, 0) !=
          hipSuccess)
    return 0;
  const long long slots = (long long)cus * per_cu;
  while (gpf > 1 && (long long)gpf * ww > slots) gpf /= 2;
  return gpf >= 2 && (long long)gpf * ww <= slots ? gpf : 0;
}

size_t mode_workspace_floats(int ww, int XX, int YY, int mm) {
  // y and bc, then (256-byte aligned) the grid kernel's partial sums:
  // kModeMaxShare rows of NQ + NPAIR doubles per frequency.
  const size_t nall = 3 * (size_t)mm + (size_t)mm * (mm + 1) / 2;
  return 2 * (size_t)ww * 2 * XX * YY * mm + 64 + (size_t)ww * kModeMaxShare * nall * 2;
}

template <int MM>
static hipError_t launch_mode_mm(ModeArgs a, int ww, hipStream_t s) {
  const int gpf = grid_share<MM>(ww, a.XX, a.YY);
  if (gpf > 0) {
    a.gpf = gpf;
    void* args[] = {&a};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_mode_grid<MM>),
                                      dim3(ww * gpf), dim3(kModeThreads), args, 0, s);
  }
  hipLaunchKernelGGL((k_mode_subspace<MM>), dim3(ww), dim3(kModeThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_mode(int ww, int XX, int YY, int mm, const float* eps, const float* omega,
                       const float* shift, float* x, float* w, float* err, int* iters,
                       int max_iters, float tol, float* ws, hipStream_t s) {
  ModeArgs a;
  a.XX = XX;
  a.YY = YY;
  a.max_iters = max_iters;
  a.tol = tol;
  a.eps = eps;
  a.omega = omega;
  a.shift = shift;
  a.x = x;
  const size_t vec = (size_t)ww * 2 * XX * YY * mm;
  a.y = ws;
  a.bc = ws + vec;
  const uintptr_t pbase = (reinterpret_cast<uintptr_t>(ws + 2 * vec) + 255) & ~uintptr_t(255);
  a.partial = reinterpret_cast<double*>(pbase);
  a.gpf = 1;
  a.w = w;
  a.err = err;
  a.iters = iters;
  switch (mm) {
    case 1: return launch_mode_mm<1>(a, ww, s);
    case 2: return launch_mode_mm<2>(a, ww, s);
    case 3: return launch_mode_mm<3>(a, ww, s);
    case 4: return launch_mode_mm<4>(a, ww, s);
    case 5: return launch_mode_mm<5>(a, ww, s);
    case 6: return launch_mode_mm<6>(a, ww, s);
    case 7: return launch_mode_mm<7>(a, ww, s);
    case 8: return launch_mode_mm<8>(a, ww, s);
    case 9: return launch_mode_mm<9>(a, ww, s);
    case 10: return launch_mode_mm<10>(a, ww, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace pjz
