// Waveguide mode solver on the device: the subspace iteration of pjz's
// mode() (src/pjz/_mode.py:99-142) over the operator of a 2-D cross-section
// (_operator, src/pjz/_mode.py:22-50), run to convergence inside one launch.
//
// One workgroup per frequency owns that frequency's whole iteration, so the
// host never synchronises inside the loop (a torch version of the same loop
// is bound by ~30 launches and one host read per iteration).  Per iteration:
//   1. x = orthonormal basis of y (CholeskyQR2: Gram matrix in double,
//      Cholesky and the triangular inverse on one lane, x = y R^-1, twice);
//   2. y = A x, the 5-point operator in two passes (the scalar fields
//      b = (d-y u0 ... ) / eps_z and c = div u, then the vector update);
//   3. w_k = <x_k, y_k>, err_k = |y_k - w_k x_k| (double sums).
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
constexpr int kChunk = 16;  // doubles per block reduction

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
  double red[kModeWaves][kChunk];
  double sum[NPAIR > NQ ? NPAIR : NQ];
  double R[MM * MM];   // Cholesky factor, upper (G = R^T R)
  double Ri[MM * MM];  // its inverse, upper
  int go;
};

// Sums acc[0..P) over the workgroup into out[0..P) (LDS); fixed order.
template <int MM>
__device__ inline void block_sum(const double (&acc)[kChunk], int P, ModeLds<MM>& L, double* out) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int p = 0; p < kChunk; ++p) {
    if (p < P) {
      double v = acc[p];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) L.red[wv][p] = v;
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < P) {
    double s = 0.0;
    for (int q = 0; q < kModeWaves; ++q) s += L.red[q][threadIdx.x];
    out[threadIdx.x] = s;
  }
  __syncthreads();
}

// Index of Gram entry (a, b), a <= b, in the row-by-row upper triangle.
template <int MM>
__device__ constexpr int tri(int a, int b) { return a * MM - a * (a - 1) / 2 + (b - a); }

// G = v^T v over the n rows of v (MM columns) into L.sum (upper triangle),
// then, on one lane, R (G = R^T R) and Ri = R^-1 in LDS.  The loops over
// modes unroll completely, so every register index is a constant.
template <int MM>
__device__ void gram_chol(const float* __restrict__ v, int n, ModeLds<MM>& L) {
  constexpr int NPAIR = ModeLds<MM>::NPAIR;
#pragma unroll
  for (int pc = 0; pc < NPAIR; pc += kChunk) {
    double acc[kChunk];
#pragma unroll
    for (int p = 0; p < kChunk; ++p) acc[p] = 0.0;
    for (int r = threadIdx.x; r < n; r += kModeThreads) {
      float row[MM];
#pragma unroll
      for (int k = 0; k < MM; ++k) row[k] = v[r * MM + k];
#pragma unroll
      for (int a = 0; a < MM; ++a) {
#pragma unroll
        for (int b = a; b < MM; ++b) {
          const int q = tri<MM>(a, b) - pc;
          if (q >= 0 && q < kChunk) acc[q] += (double)row[a] * (double)row[b];
        }
      }
    }
    block_sum<MM>(acc, min(kChunk, NPAIR - pc), L, L.sum + pc);
  }
  if (threadIdx.x == 0) {
    double* R = L.R;
    double* Ri = L.Ri;
    for (int i = 0; i < MM * MM; ++i) R[i] = Ri[i] = 0.0;
    for (int j = 0; j < MM; ++j) {
      double d = L.sum[tri<MM>(j, j)];
      for (int k = 0; k < j; ++k) d -= R[k * MM + j] * R[k * MM + j];
      // A start vector inside the span of the others (degenerate input
      // only) keeps a tiny pivot instead of a NaN.
      R[j * MM + j] = sqrt(d > 1e-300 ? d : 1e-300);
      for (int i = j + 1; i < MM; ++i) {
        double t = L.sum[tri<MM>(j, i)];
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
  __syncthreads();
}

// dst = src Ri, row by row (dst may alias src).
template <int MM>
__device__ void apply_ri(const float* src, float* dst, int n, const ModeLds<MM>& L) {
  for (int r = threadIdx.x; r < n; r += kModeThreads) {
    float row[MM];
#pragma unroll
    for (int k = 0; k < MM; ++k) row[k] = src[r * MM + k];
#pragma unroll
    for (int k = 0; k < MM; ++k) {
      double t = 0.0;
#pragma unroll
      for (int j = 0; j <= k; ++j) t += (double)row[j] * L.Ri[j * MM + k];
      dst[r * MM + k] = (float)t;
    }
  }
  __syncthreads();
}

// y = A x for one frequency (om2 = omega^2, s = shift): _operator of
// src/pjz/_mode.py:22-50, with the same grouping of the float operations.
template <int MM>
__device__ void apply_op(const float* __restrict__ x, float* __restrict__ y, float* __restrict__ bc,
                         const float* __restrict__ eps, int XX, int YY, float om2, float s) {
  const int P = XX * YY, cells = P * MM;
  const float* ex = eps;
  const float* ey = eps + P;
  const float* ez = eps + 2 * P;
  for (int e = threadIdx.x; e < cells; e += kModeThreads) {
    const int p = e / MM, k = e - p * MM;
    const int i = p / YY, j = p - i * YY;
    const int im = i == 0 ? XX - 1 : i - 1, ip = i == XX - 1 ? 0 : i + 1;
    const int jm = j == 0 ? YY - 1 : j - 1, jp = j == YY - 1 ? 0 : j + 1;
    const float* x0 = x;
    const float* x1 = x + cells;
    const float u0 = x0[e], u1 = x1[e];
    const float b = (-(u0 - x0[(i * YY + jm) * MM + k]) + (u1 - x1[(im * YY + j) * MM + k])) / ez[p];
    const float c = (x0[(ip * YY + j) * MM + k] - u0) + (x1[(i * YY + jp) * MM + k] - u1);
    bc[e] = b;
    bc[cells + e] = c;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < cells; e += kModeThreads) {
    const int p = e / MM, k = e - p * MM;
    const int i = p / YY, j = p - i * YY;
    const int im = i == 0 ? XX - 1 : i - 1, ip = i == XX - 1 ? 0 : i + 1;
    const int jm = j == 0 ? YY - 1 : j - 1, jp = j == YY - 1 ? 0 : j + 1;
    const float* bb = bc;
    const float* cc = bc + cells;
    const float b = bb[e], c = cc[e];
    const float u0 = x[e], u1 = x[cells + e];
    const float a0 = (om2 * ey[p] - s) * u0, a1 = (om2 * ex[p] - s) * u1;
    const float b0 = -(bb[(i * YY + jp) * MM + k] - b) * ey[p];
    const float b1 = (bb[(ip * YY + j) * MM + k] - b) * ex[p];
    y[e] = (a0 + b0) + (c - cc[(im * YY + j) * MM + k]);
    y[cells + e] = (a1 + b1) + (c - cc[(i * YY + jm) * MM + k]);
  }
  __syncthreads();
}

template <int MM>
__global__ __launch_bounds__(kModeThreads) void k_mode_subspace(ModeArgs a) {
  __shared__ ModeLds<MM> L;
  const int f = blockIdx.x;
  const int n = 2 * a.XX * a.YY;  // rows of the (n, MM) matrices
  const long long off = (long long)f * n * MM;
  float* x = a.x + off;
  float* y = a.y + off;
  float* bc = a.bc + off;
  const float om = a.omega[f];
  const float om2 = om * om;
  const float s = a.shift[f];
  apply_op<MM>(x, y, bc, a.eps, a.XX, a.YY, om2, s);
  constexpr int NQ = ModeLds<MM>::NQ;
  int it = 0;
  bool go = true;  // max err > tol; err starts at +inf
  while (it < a.max_iters && go) {
    gram_chol<MM>(y, n, L);
    apply_ri<MM>(y, x, n, L);
    gram_chol<MM>(x, n, L);
    apply_ri<MM>(x, x, n, L);
    apply_op<MM>(x, y, bc, a.eps, a.XX, a.YY, om2, s);
    double axy[MM], ayy[MM], axx[MM];
#pragma unroll
    for (int k = 0; k < MM; ++k) axy[k] = ayy[k] = axx[k] = 0.0;
    for (int r = threadIdx.x; r < n; r += kModeThreads) {
#pragma unroll
      for (int k = 0; k < MM; ++k) {
        const double xv = x[r * MM + k], yv = y[r * MM + k];
        axy[k] += xv * yv;
        ayy[k] += yv * yv;
        axx[k] += xv * xv;
      }
    }
#pragma unroll
    for (int qc = 0; qc < NQ; qc += kChunk) {
      double acc[kChunk];
#pragma unroll
      for (int p = 0; p < kChunk; ++p) {
        const int q = qc + p;
        acc[p] = q < MM ? axy[q < MM ? q : 0]
                 : q < 2 * MM ? ayy[q < 2 * MM ? q - MM : 0]
                 : q < NQ ? axx[q < NQ ? q - 2 * MM : 0] : 0.0;
      }
      block_sum<MM>(acc, min(kChunk, NQ - qc), L, L.sum + qc);
    }
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
    }
    __syncthreads();
    go = L.go;
    ++it;
    __syncthreads();
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
