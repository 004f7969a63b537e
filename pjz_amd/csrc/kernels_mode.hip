// Waveguide mode solver on the device: the subspace iteration of pjz's
// mode() (src/pjz/_mode.py:99-142) over the operator of a 2-D cross-section
// (_operator, src/pjz/_mode.py:22-50), run to convergence inside one launch.
//
// One workgroup per frequency owns that frequency's whole iteration, so the
// host never synchronises inside the loop (a torch version of the same loop
// is bound by ~30 launches and one host read per iteration).  Per iteration,
// four passes over the vectors:
//   1. x = y R1^-1, summing the Gram matrix of x as the rows are written;
//   2. x = x R2^-1 (1 and 2 are CholeskyQR2: Gram matrices in double,
//      Cholesky and the triangular inverse on one lane);
//   3. y = A x, first stencil pass: the scalar fields b = (-(d-y u0) +
//      d-x u1) / eps_z and c = d+x u0 + d+y u1;
//   4. second stencil pass, the vector update, summing w_k = <x_k, y_k>,
//      |y_k|^2, |x_k|^2 (so err_k = |y_k - w_k x_k|) and the Gram matrix of
//      y, which gives the next iteration's R1.
// The loop stops after max_iters or once every err_k <= tol, the test pjz
// makes before each iteration.  Vectors are (2, XX, YY, MM) with the MM modes
// innermost, as pjz lays them out; the periodic differences wrap in x and y
// (jnp.roll in pjz).  Every pass is a stride loop over the cells, so any
// cross-section size works; fields live in global memory (L2-resident at the
// usual sizes) and the reductions go through LDS.
#include "fdtd_common.h"

namespace pjz {

constexpr int kModeThreads = 512;
constexpr int kModeWaves = kModeThreads / 64;

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

// acc[tri(a, b)] += v[a] v[b] for a <= b (the loops unroll completely, so
// every register index is a constant).
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

// First stencil pass of y = A x (_operator, src/pjz/_mode.py:22-50): the
// scalar fields b = (-(d-y u0) + d-x u1) / eps_z (backward differences) and
// c = d+x u0 + d+y u1 (forward differences), periodic in x and y.
template <int MM>
__device__ __forceinline__ void op_pass1(const float* __restrict__ x, float* __restrict__ bc,
                         const float* __restrict__ eps, int XX, int YY) {
  const int P = XX * YY, cells = P * MM;
  const float* ez = eps + 2 * P;
  for (int e = threadIdx.x; e < cells; e += kModeThreads) {
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
  __syncthreads();
}

// Whether the second stencil pass also sums the Gram matrix of y: its
// 3 MM + MM (MM + 1) / 2 double accumulators fit in registers up to 6 modes;
// above, a separate pass sums the Gram matrix.
template <int MM>
constexpr bool kFuseGram = MM <= 6;

// Second stencil pass, one cell (both components, all modes) per thread, with
// the same grouping of the float operations as _operator: y = A x, and into
// acc the Rayleigh sums (x.y, y.y, x.x per mode) and, when P has room for it,
// the Gram matrix of y.
template <int MM, int P>
__device__ __forceinline__ void op_pass2(const float* __restrict__ x, float* __restrict__ y,
                         const float* __restrict__ bc, const float* __restrict__ eps, int XX,
                         int YY, float om2, float s, double (&acc)[P]) {
  const int PC = XX * YY, cells = PC * MM;
  const float* ex = eps;
  const float* ey = eps + PC;
  const float* bb = bc;
  const float* cc = bc + cells;
  for (int p = threadIdx.x; p < PC; p += kModeThreads) {
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

// The stencil's second pass with all its sums into L.sum (Rayleigh sums, then
// the Gram matrix of y), in one pass or two.
template <int MM>
__device__ __forceinline__ void op_pass2_sums(const float* __restrict__ x, float* __restrict__ y,
                              const float* __restrict__ bc, const float* __restrict__ eps,
                              int XX, int YY, float om2, float s, int n, ModeLds<MM>& L) {
  constexpr int NQ = ModeLds<MM>::NQ, NPAIR = ModeLds<MM>::NPAIR;
  if constexpr (kFuseGram<MM>) {
    double acc[NQ + NPAIR];
#pragma unroll
    for (int q = 0; q < NQ + NPAIR; ++q) acc[q] = 0.0;
    op_pass2<MM>(x, y, bc, eps, XX, YY, om2, s, acc);
    block_sum<MM>(acc, L, L.sum);
  } else {
    {
      double acc[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
      op_pass2<MM>(x, y, bc, eps, XX, YY, om2, s, acc);
      block_sum<MM>(acc, L, L.sum);
    }
    double acc[NPAIR];
#pragma unroll
    for (int q = 0; q < NPAIR; ++q) acc[q] = 0.0;
    for (int r = threadIdx.x; r < n; r += kModeThreads) {
      float v[MM];
#pragma unroll
      for (int k = 0; k < MM; ++k) v[k] = y[r * MM + k];
      add_gram<MM>(acc, 0, v);
    }
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
  // y = A x0 and its Gram matrix; then R^-1 for the first QR.
  op_pass1<MM>(x, bc, a.eps, a.XX, a.YY);
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
      for (int r = threadIdx.x; r < n; r += kModeThreads) {
        float v[MM];
#pragma unroll
        for (int k = 0; k < MM; ++k) v[k] = y[r * MM + k];
        row_ri<MM>(v, L);
#pragma unroll
        for (int k = 0; k < MM; ++k) x[r * MM + k] = v[k];
        if constexpr (MM <= 8) add_gram<MM>(acc, 0, v);
      }
      if constexpr (MM > 8) {
        // Above 8 modes the Gram sum takes a pass of its own (registers).
        __syncthreads();
        for (int r = threadIdx.x; r < n; r += kModeThreads) {
          float v[MM];
#pragma unroll
          for (int k = 0; k < MM; ++k) v[k] = x[r * MM + k];
          add_gram<MM>(acc, 0, v);
        }
      }
      block_sum<MM>(acc, L, L.sum + NQ);
      if (threadIdx.x == 0) chol_inv<MM>(L.sum + NQ, L);
      __syncthreads();
      for (int r = threadIdx.x; r < n; r += kModeThreads) {
        float v[MM];
#pragma unroll
        for (int k = 0; k < MM; ++k) v[k] = x[r * MM + k];
        row_ri<MM>(v, L);
#pragma unroll
        for (int k = 0; k < MM; ++k) x[r * MM + k] = v[k];
      }
      __syncthreads();
    }
    // y = A x with the Rayleigh sums and the Gram matrix of the new y.
    op_pass1<MM>(x, bc, a.eps, a.XX, a.YY);
    op_pass2_sums<MM>(x, y, bc, a.eps, a.XX, a.YY, om2, s, n, L);
    if (threadIdx.x == 0) {
      double emax = 0.0;
      bool nan = false;
      for (int k = 0; k < MM; ++k) {
        const double xy = L.sum[k], yy = L.sum[MM + k], xx = L.sum[2 * MM + k];
        const double e2 = yy - 2.0 * xy * xy + xy * xy * xx;
        const bool bad = !(e2 == e2);
        const double e = sqrt(e2 > 0.0 ? e2 : 0.0);
        nan |= bad;
        emax = e > emax ? e : emax;
        a.w[f * MM + k] = (float)xy;
        a.err[f * MM + k] = bad ? __builtin_nanf("") : (float)e;
      }
      // pjz continues while max(err) > tol, which a NaN ends.
      L.go = !nan && emax > (double)a.tol;
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

size_t mode_workspace_floats(int ww, int XX, int YY, int mm) {
  return 2 * (size_t)ww * 2 * XX * YY * mm;
}

template <int MM>
static void launch_mode_mm(const ModeArgs& a, int ww, hipStream_t s) {
  hipLaunchKernelGGL((k_mode_subspace<MM>), dim3(ww), dim3(kModeThreads), 0, s, a);
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
  a.y = ws;
  a.bc = ws + (size_t)ww * 2 * XX * YY * mm;
  a.w = w;
  a.err = err;
  a.iters = iters;
  switch (mm) {
    case 1: launch_mode_mm<1>(a, ww, s); break;
    case 2: launch_mode_mm<2>(a, ww, s); break;
    case 3: launch_mode_mm<3>(a, ww, s); break;
    case 4: launch_mode_mm<4>(a, ww, s); break;
    case 5: launch_mode_mm<5>(a, ww, s); break;
    case 6: launch_mode_mm<6>(a, ww, s); break;
    case 7: launch_mode_mm<7>(a, ww, s); break;
    case 8: launch_mode_mm<8>(a, ww, s); break;
    case 9: launch_mode_mm<9>(a, ww, s); break;
    case 10: launch_mode_mm<10>(a, ww, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace pjz
