// Set-up kernels (coefficients, state clearing) and the reference stepper:
// two launches per time step (H half step, then E half step), one thread per
// cell.  The reference stepper is the simplest correct statement of the update
// on the GPU; the streaming kernel (kernels_stream.hip) is the fast path.
#include "fdtd_common.h"

namespace pjz {

// ---------------------------------------------------------------- set-up
// cb = (dt/eps)/(1+q), q = sigma_abs*dt/(2 eps), eps edge-replicated outside
// the epsilon window.  Computed in double, stored as float.
__global__ void k_prepare_cb(const float* __restrict__ eps, const float* __restrict__ mask,
                             float* __restrict__ cb, Geom g, int x0, int y0, int z0,
                             int ex, int ey, int ez, double dt) {
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= 3 * g.N) return;
  int k = (int)(idx % g.Z);
  long long r = idx / g.Z;
  int j = (int)(r % g.Y);
  r /= g.Y;
  int i = (int)(r % g.X);
  int c = (int)(r / g.X);
  int ii = min(max(i - x0, 0), ex - 1);
  int jj = min(max(j - y0, 0), ey - 1);
  int kk = min(max(k - z0, 0), ez - 1);
  double e = eps[(((long long)c * ex + ii) * ey + jj) * ez + kk];
  double s = mask[((long long)c * g.X + i) * g.Y + j];
  double q = s * dt / (2.0 * e);
  cb[idx] = (float)((dt / e) / (1.0 + q));
}

// Per-z CPML table from the (zz, 2) kappa/sigma/alpha arrays.
__global__ void k_prepare_zt(const float* __restrict__ kappa, const float* __restrict__ sigma,
                             const float* __restrict__ alpha, float* __restrict__ zt, int Z,
                             double dt) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= Z) return;
  for (int col = 0; col < 2; ++col) {
    double kap = kappa[2 * k + col], sig = sigma[2 * k + col], alp = alpha[2 * k + col];
    double ik = isinf(kap) ? 0.0 : 1.0 / kap;
    double b = exp(-(sig * ik + alp) * dt);
    double den = sig * ik + alp;
    double c = (sig == 0.0 || ik == 0.0 || den == 0.0) ? 0.0 : sig * ik * ik * (b - 1.0) / den;
    int base = col == 0 ? IKE : IKH;
    zt[(base + 0) * Z + k] = (float)ik;
    zt[(base + 1) * Z + k] = (float)b;
    zt[(base + 2) * Z + k] = (float)c;
  }
}

__global__ void k_fill_zero(float* __restrict__ p, long long n) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0.0f;
}

static inline unsigned blocks_for(long long n, int bs) {
  long long b = (n + bs - 1) / bs;
  return (unsigned)(b < 1 ? 1 : b);
}

hipError_t launch_prepare(const Geom& g, const pjz_fdtd_problem& p, const float* eps,
                          const float* mask, const float* kappa, const float* sigma,
                          const float* alpha, float* cb, float* zt, hipStream_t s) {
  hipLaunchKernelGGL(k_prepare_cb, dim3(blocks_for(3 * g.N, 256)), dim3(256), 0, s, eps, mask,
                     cb, g, p.x0, p.y0, p.z0, p.ex, p.ey, p.ez, (double)p.dt);
  hipLaunchKernelGGL(k_prepare_zt, dim3(blocks_for(g.Z, 64)), dim3(64), 0, s, kappa, sigma,
                     alpha, zt, g.Z, (double)p.dt);
  return hipGetLastError();
}

// Streaming copy of n16 16-byte words with nontemporal loads and stores: the
// device-copy ceiling the steppers' HBM traffic is compared against.
typedef float f32x4v __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(1024) void k_copy_stream(const f32x4v* __restrict__ src,
                                                      f32x4v* __restrict__ dst, long long n16) {
  long long i = blockIdx.x * 1024LL + threadIdx.x;
  const long long stride = (long long)gridDim.x * 1024;
  for (; i < n16; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

hipError_t launch_copy_stream(void* dst, const void* src, size_t n16, hipStream_t s) {
  unsigned nb = blocks_for((long long)n16, 1024);
  if (nb > 16384u) nb = 16384u;
  hipLaunchKernelGGL(k_copy_stream, dim3(nb), dim3(1024), 0, s, (const f32x4v*)src, (f32x4v*)dst,
                     (long long)n16);
  return hipGetLastError();
}

hipError_t launch_zero(float* ptr, size_t n, hipStream_t s) {
  unsigned nb = blocks_for((long long)n, 256);
  if (nb > 65536u) nb = 65536u;
  hipLaunchKernelGGL(k_fill_zero, dim3(nb), dim3(256), 0, s, ptr, (long long)n);
  return hipGetLastError();
}

// ---------------------------------------------------------------- reference stepper
__device__ inline int pml_index(int k, const Geom& g) {
  return k < g.lo ? k : k - (g.Z - g.hi) + g.lo;
}

__global__ void k_ref_h(const float* __restrict__ E, float* __restrict__ H,
                        float* __restrict__ psi, const float* __restrict__ zt, Geom g, float dt) {
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= g.N) return;
  const long long N = g.N, YZ = (long long)g.Y * g.Z;
  int k = (int)(idx % g.Z);
  long long xy = idx / g.Z;
  int j = (int)(xy % g.Y);
  int i = (int)(xy / g.Y);
  const float *Ex = E, *Ey = E + N, *Ez = E + 2 * N;
  float ex = Ex[idx], ey = Ey[idx], ez = Ez[idx];
  float ex_zp = k + 1 < g.Z ? Ex[idx + 1] : 0.0f;
  float ey_zp = k + 1 < g.Z ? Ey[idx + 1] : 0.0f;
  float ex_yp = j + 1 < g.Y ? Ex[idx + g.Z] : 0.0f;
  float ez_yp = j + 1 < g.Y ? Ez[idx + g.Z] : 0.0f;
  float ey_xp = i + 1 < g.X ? Ey[idx + YZ] : 0.0f;
  float ez_xp = i + 1 < g.X ? Ez[idx + YZ] : 0.0f;
  float dzEy = ey_zp - ey, dzEx = ex_zp - ex;
  float ik = zt[IKH * g.Z + k];
  float tEy = ik * dzEy, tEx = ik * dzEx;
  if (k < g.lo || k >= g.Z - g.hi) {
    long long pstride = (long long)g.X * g.Y * g.NP;
    long long pi = xy * g.NP + pml_index(k, g);
    float b = zt[BH * g.Z + k], c = zt[CH * g.Z + k];
    float p0 = b * psi[PSI_HX * pstride + pi] + c * dzEy;
    float p1 = b * psi[PSI_HY * pstride + pi] + c * dzEx;
    psi[PSI_HX * pstride + pi] = p0;
    psi[PSI_HY * pstride + pi] = p1;
    tEy += p0;
    tEx += p1;
  }
  H[idx] = H[idx] - dt * ((ez_yp - ez) - tEy);
  H[N + idx] = H[N + idx] - dt * (tEx - (ez_xp - ez));
  H[2 * N + idx] = H[2 * N + idx] - dt * ((ey_xp - ey) - (ex_yp - ex));
}

__global__ void k_ref_e(float* __restrict__ E, const float* __restrict__ H,
                        float* __restrict__ psi, const float* __restrict__ zt,
                        const float* __restrict__ cb, const float* __restrict__ mask, Geom g,
                        Src src, int n, OutWin win, int out_idx, float* __restrict__ Ecap) {
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= g.N) return;
  const long long N = g.N, YZ = (long long)g.Y * g.Z;
  int k = (int)(idx % g.Z);
  long long xy = idx / g.Z;
  int j = (int)(xy % g.Y);
  int i = (int)(xy / g.Y);
  const float *Hx = H, *Hy = H + N, *Hz = H + 2 * N;
  float hx = Hx[idx], hy = Hy[idx], hz = Hz[idx];
  float hx_zm = k > 0 ? Hx[idx - 1] : 0.0f;
  float hy_zm = k > 0 ? Hy[idx - 1] : 0.0f;
  float hx_ym = j > 0 ? Hx[idx - g.Z] : 0.0f;
  float hz_ym = j > 0 ? Hz[idx - g.Z] : 0.0f;
  float hy_xm = i > 0 ? Hy[idx - YZ] : 0.0f;
  float hz_xm = i > 0 ? Hz[idx - YZ] : 0.0f;
  float dzHy = hy - hy_zm, dzHx = hx - hx_zm;
  float ik = zt[IKE * g.Z + k];
  float tHy = ik * dzHy, tHx = ik * dzHx;
  if (k < g.lo || k >= g.Z - g.hi) {
    long long pstride = (long long)g.X * g.Y * g.NP;
    long long pi = xy * g.NP + pml_index(k, g);
    float b = zt[BE * g.Z + k], c = zt[CE * g.Z + k];
    float p0 = b * psi[PSI_EX * pstride + pi] + c * dzHy;
    float p1 = b * psi[PSI_EY * pstride + pi] + c * dzHx;
    psi[PSI_EX * pstride + pi] = p0;
    psi[PSI_EY * pstride + pi] = p1;
    tHy += p0;
    tHx += p1;
  }
  float cx = (hz - hz_ym) - tHy;
  float cy = tHx - (hz - hz_xm);
  float cz = (hy - hy_xm) - (hx - hx_ym);
  // Source current J(n) on its plane(s).
  if (src.axis == 0) {
    if (i == src.pos || i == src.pos - 1) {
      float w = src.wave[2 * n + (i == src.pos ? 0 : 1)];
      long long o = (long long)j * g.Z + k;
      cy += src.field[o] * w;
      cz += src.field[YZ + o] * w;
    }
  } else if (src.axis == 1) {
    if (j == src.pos || j == src.pos - 1) {
      float w = src.wave[2 * n + (j == src.pos ? 0 : 1)];
      long long XZ = (long long)g.X * g.Z;
      long long o = (long long)i * g.Z + k;
      cx += src.field[o] * w;
      cz += src.field[XZ + o] * w;
    }
  } else if (k == src.pos) {
    long long XY = (long long)g.X * g.Y;
    float jx = 0.0f, jy = 0.0f;
    for (int q = 0; q < src.quad; ++q) {
      float w = src.wave[2 * n + q];
      jx += src.field[(2 * q + 0) * XY + xy] * w;
      jy += src.field[(2 * q + 1) * XY + xy] * w;
    }
    cx += jx;
    cy += jy;
  }
  const float c3[3] = {cx, cy, cz};
  bool in_win = out_idx >= 0 && i >= win.x0 && i < win.x0 + win.ex && j >= win.y0 &&
                j < win.y0 + win.ey && k >= win.z0 && k < win.z0 + win.ez;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float b = cb[c * N + idx];
    float s = mask[((long long)c * g.X + i) * g.Y + j];
    float a = 1.0f - s * b;
    const float eold = E[c * N + idx];
    float e = a * eold + b * c3[c];
    E[c * N + idx] = e;
    if (Ecap) Ecap[c * N + idx] = eold - e;
    if (in_win) {
      long long o = ((((long long)out_idx * 3 + c) * win.ex + (i - win.x0)) * win.ey +
                     (j - win.y0)) * win.ez + (k - win.z0);
      win.out[o] = e;
    }
  }
}

hipError_t run_reference(const Geom& g, const pjz_fdtd_problem& p, const float* mask,
                         const Src& src, const OutWin& win, char* ws, const Layout& L,
                         const StepRange& r, hipStream_t s) {
  float* E = (float*)(ws + L.E0);
  float* H = (float*)(ws + L.H0);
  float* cb = (float*)(ws + L.cb);
  float* psi = (float*)(ws + L.psi);
  const float* zt = (const float*)(ws + L.zt);
  unsigned nb = blocks_for(g.N, 256);
  for (int n = r.begin; n < r.end; ++n) {
    int out_idx = -1;
    if (win.out && n >= p.out_start && n < p.out_stop && (n - p.out_start) % p.out_step == 0)
      out_idx = (n - p.out_start) / p.out_step;
    float* cap = r.capture ? r.capture + (long long)(n - r.begin) * 3 * g.N : nullptr;
    hipLaunchKernelGGL(k_ref_h, dim3(nb), dim3(256), 0, s, E, H, psi, zt, g, p.dt);
    hipLaunchKernelGGL(k_ref_e, dim3(nb), dim3(256), 0, s, E, H, psi, zt, cb, mask, g, src, n,
                       win, out_idx, cap);
  }
  *r.cur = 0;
  return hipGetLastError();
}

__global__ void k_copy(float* __restrict__ dst, const float* __restrict__ src, long long n) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}

hipError_t launch_copy(float* dst, const float* src, size_t n, hipStream_t s) {
  unsigned nb = blocks_for((long long)n, 256);
  if (nb > 65536u) nb = 65536u;
  hipLaunchKernelGGL(k_copy, dim3(nb), dim3(256), 0, s, dst, src, (long long)n);
  return hipGetLastError();
}

}  // namespace pjz
