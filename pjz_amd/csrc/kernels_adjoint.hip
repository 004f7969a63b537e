// Discrete adjoint of the update: reverse time steps, transposed curls,
// transposed CPML recursion and absorber, gradient with respect to eps.
//
// Forward step n (E' = E^(n+1), H' = H^(n+1/2)):
//   psiH' = bH psiH + cH Dzf E;   H' = H - dt Cf E          (z terms ikH Dzf E + psiH')
//   psiE' = bE psiE + cE Dzb H';  E' = ca E + cb K,  K = Cb H' - J  (z terms ikE Dzb H' + psiE')
// Reverse step n, with lK = cb lE':
//   k_adj_h (H cells): lH <- lH + Cb^T lK - Dzf(cE (lpsiE' -+ lK)),  u <- bH u +- dt lH
//   k_adj_e (E cells): lE <- ca lE' + dt Cf^T lH - Dzb(cH u),  lpsiE <- bE (lpsiE' -+ lK),
//                      G += lE' (E^n - E^(n+1))
// u is the psi_H cotangent before the bH scaling that the next reverse step
// applies, so no kernel reads a neighbour's psi value another thread writes.
// Gradient: dL/dca = sum lE' E^n = A and dL/dcb = sum lE' K = (B - ca A)/cb
// with B = sum lE' E^(n+1) (since cb K = E' - ca E).  With ca = 1 - s cb and
// cb = dt / (eps + s dt / 2): dca/deps = s cb^2/dt, dcb/deps = -cb^2/dt, so
// dL/deps = (cb/dt) (A - B) = (cb/dt) G.  Accumulating G from the per-step
// change of E halves the accumulator traffic and avoids the cancellation of
// A - B.
#include "fdtd_common.h"

namespace pjz {

static inline unsigned nblocks(long long n, int bs) {
  long long b = (n + bs - 1) / bs;
  return (unsigned)(b < 1 ? 1 : b);
}

__device__ inline bool in_pml(int k, const Geom& g) { return k < g.lo || k >= g.Z - g.hi; }
__device__ inline int pml_at(int k, const Geom& g) {
  return k < g.lo ? k : k - (g.Z - g.hi) + g.lo;
}

struct AdjArgs {
  float* lE;
  float* lH;
  float* lu;   // (2, X, Y, NP): psi_H cotangent pair (x, y)
  float* lpE;  // (2, X, Y, NP): psi_E cotangent pair (x, y)
  float* G;
  const float* cb;
  const float* mask;
  const float* zt;
  const float* D;  // E^n - E^(n+1) of the step being reversed
  Geom g;
  float dt;
};

__device__ inline void adj_h_cell(const AdjArgs& a, long long idx) {
  const Geom& g = a.g;
  const long long N = g.N, YZ = (long long)g.Y * g.Z, P = (long long)g.X * g.Y * g.NP;
  const int k = (int)(idx % g.Z);
  const long long xy = idx / g.Z;
  const int j = (int)(xy % g.Y);
  const int i = (int)(xy / g.Y);
  const float* cb = a.cb;
  const float* lE = a.lE;
  auto lK = [&](int c, long long o) { return cb[c * N + o] * lE[c * N + o]; };
  const float kx = lK(0, idx), ky = lK(1, idx), kz = lK(2, idx);
  const bool zp = k + 1 < g.Z, yp = j + 1 < g.Y, xp = i + 1 < g.X;
  const float kx_zp = zp ? lK(0, idx + 1) : 0.0f;
  const float ky_zp = zp ? lK(1, idx + 1) : 0.0f;
  const float kz_yp = yp ? lK(2, idx + g.Z) : 0.0f;
  const float kx_yp = yp ? lK(0, idx + g.Z) : 0.0f;
  const float kz_xp = xp ? lK(2, idx + YZ) : 0.0f;
  const float ky_xp = xp ? lK(1, idx + YZ) : 0.0f;
  const float ikE = a.zt[IKE * g.Z + k];
  const float ikE_p = zp ? a.zt[IKE * g.Z + k + 1] : 0.0f;
  // cE (lpsiE' -+ lK) at k and k+1 (zero outside the PML)
  float vx = 0.0f, vy = 0.0f, vx_p = 0.0f, vy_p = 0.0f;
  if (in_pml(k, g)) {
    const long long pi = xy * g.NP + pml_at(k, g);
    const float c = a.zt[CE * g.Z + k];
    vx = c * (a.lpE[pi] - kx);
    vy = c * (a.lpE[P + pi] + ky);
  }
  if (zp && in_pml(k + 1, g)) {
    const long long pi = xy * g.NP + pml_at(k + 1, g);
    const float c = a.zt[CE * g.Z + k + 1];
    vx_p = c * (a.lpE[pi] - kx_zp);
    vy_p = c * (a.lpE[P + pi] + ky_zp);
  }
  const float hx = a.lH[idx] - (ikE_p * ky_zp - ikE * ky) + (kz_yp - kz) - (vy_p - vy);
  const float hy = a.lH[N + idx] + (ikE_p * kx_zp - ikE * kx) - (kz_xp - kz) - (vx_p - vx);
  const float hz = a.lH[2 * N + idx] - (kx_yp - kx) + (ky_xp - ky);
  a.lH[idx] = hx;
  a.lH[N + idx] = hy;
  a.lH[2 * N + idx] = hz;
  if (in_pml(k, g)) {
    const long long pi = xy * g.NP + pml_at(k, g);
    const float b = a.zt[BH * g.Z + k];
    a.lu[pi] = b * a.lu[pi] + a.dt * hx;
    a.lu[P + pi] = b * a.lu[P + pi] - a.dt * hy;
  }
}

__device__ inline void adj_e_cell(const AdjArgs& a, long long idx, const float* D) {
  const Geom& g = a.g;
  const long long N = g.N, YZ = (long long)g.Y * g.Z, P = (long long)g.X * g.Y * g.NP;
  const int k = (int)(idx % g.Z);
  const long long xy = idx / g.Z;
  const int j = (int)(xy % g.Y);
  const int i = (int)(xy / g.Y);
  const float* lH = a.lH;
  const bool zm = k > 0, ym = j > 0, xm = i > 0;
  const float hx = lH[idx], hy = lH[N + idx], hz = lH[2 * N + idx];
  const float hx_zm = zm ? lH[idx - 1] : 0.0f;
  const float hy_zm = zm ? lH[N + idx - 1] : 0.0f;
  const float hx_ym = ym ? lH[idx - g.Z] : 0.0f;
  const float hz_ym = ym ? lH[2 * N + idx - g.Z] : 0.0f;
  const float hy_xm = xm ? lH[N + idx - YZ] : 0.0f;
  const float hz_xm = xm ? lH[2 * N + idx - YZ] : 0.0f;
  const float ikH = a.zt[IKH * g.Z + k];
  const float ikH_m = zm ? a.zt[IKH * g.Z + k - 1] : 0.0f;
  // cH u at k and k-1 (zero outside the PML)
  float ux = 0.0f, uy = 0.0f, ux_m = 0.0f, uy_m = 0.0f;
  if (in_pml(k, g)) {
    const long long pi = xy * g.NP + pml_at(k, g);
    const float c = a.zt[CH * g.Z + k];
    ux = c * a.lu[pi];
    uy = c * a.lu[P + pi];
  }
  if (zm && in_pml(k - 1, g)) {
    const long long pi = xy * g.NP + pml_at(k - 1, g);
    const float c = a.zt[CH * g.Z + k - 1];
    ux_m = c * a.lu[pi];
    uy_m = c * a.lu[P + pi];
  }
  const float dt = a.dt;
  const float gx = dt * ((ikH * hy - ikH_m * hy_zm) - (hz - hz_ym)) - (uy - uy_m);
  const float gy = dt * (-(ikH * hx - ikH_m * hx_zm) + (hz - hz_xm)) - (ux - ux_m);
  const float gz = dt * ((hx - hx_ym) - (hy - hy_xm));
  const float gsum[3] = {gx, gy, gz};
  float lk[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const long long o = c * N + idx;
    const float le = a.lE[o];
    const float b = a.cb[o];
    const float s = a.mask[((long long)c * g.X + i) * g.Y + j];
    a.G[o] += le * D[o];
    lk[c] = b * le;
    a.lE[o] = (1.0f - s * b) * le + gsum[c];
  }
  if (in_pml(k, g)) {
    const long long pi = xy * g.NP + pml_at(k, g);
    const float b = a.zt[BE * g.Z + k];
    a.lpE[pi] = b * (a.lpE[pi] - lk[0]);
    a.lpE[P + pi] = b * (a.lpE[P + pi] + lk[1]);
  }
}

__global__ void k_adj_h(AdjArgs a) {
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx < a.g.N) adj_h_cell(a, idx);
}

__global__ void k_adj_e(AdjArgs a) {
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx < a.g.N) adj_e_cell(a, idx, a.D);
}

static AdjArgs adj_args(const Geom& g, const pjz_fdtd_problem& p, const float* mask, char* ws,
                        const Layout& L, int acur) {
  AdjArgs a;
  a.lE = (float*)(ws + L.lE_[acur]);
  a.lH = (float*)(ws + L.lH_[acur]);
  a.lu = (float*)(ws + L.lu_[acur]);
  a.lpE = (float*)(ws + L.lpE_[acur]);
  a.G = (float*)(ws + L.accG);
  a.cb = (const float*)(ws + L.cb);
  a.zt = (const float*)(ws + L.zt);
  a.mask = mask;
  a.D = nullptr;
  a.g = g;
  a.dt = p.dt;
  return a;
}

hipError_t launch_adjoint_step(const Geom& g, const pjz_fdtd_problem& p, const float* mask,
                               char* ws, const Layout& L, const float* D, int acur,
                               hipStream_t s) {
  AdjArgs a = adj_args(g, p, mask, ws, L, acur);
  a.D = D;
  const unsigned nb = nblocks(g.N, 256);
  hipLaunchKernelGGL(k_adj_h, dim3(nb), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_adj_e, dim3(nb), dim3(256), 0, s, a);
  return hipGetLastError();
}

__global__ void k_inject(float* __restrict__ lE, const float* __restrict__ gout, Geom g, int x0,
                         int y0, int z0, int ex, int ey, int ez) {
  long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long n = 3LL * ex * ey * ez;
  if (t >= n) return;
  const int z = (int)(t % ez);
  long long r = t / ez;
  const int y = (int)(r % ey);
  r /= ey;
  const int x = (int)(r % ex);
  const int c = (int)(r / ex);
  lE[c * g.N + (((long long)(x + x0) * g.Y + (y + y0)) * g.Z + (z + z0))] += gout[t];
}

hipError_t launch_inject(const Geom& g, const pjz_fdtd_problem& p, float* lE, const float* g_out,
                         hipStream_t s) {
  const long long n = 3LL * p.ex * p.ey * p.ez;
  hipLaunchKernelGGL(k_inject, dim3(nblocks(n, 256)), dim3(256), 0, s, lE, g_out, g, p.x0, p.y0,
                     p.z0, p.ex, p.ey, p.ez);
  return hipGetLastError();
}

__global__ void k_grad_eps(const float* __restrict__ eps, const float* __restrict__ mask,
                           const float* __restrict__ G, float* __restrict__ grad, Geom g, int x0,
                           int y0, int z0, int ex, int ey, int ez, double dt) {
  long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= 3 * g.N) return;
  const int k = (int)(t % g.Z);
  long long r = t / g.Z;
  const int j = (int)(r % g.Y);
  r /= g.Y;
  const int i = (int)(r % g.X);
  const int c = (int)(r / g.X);
  const int ii = min(max(i - x0, 0), ex - 1);
  const int jj = min(max(j - y0, 0), ey - 1);
  const int kk = min(max(k - z0, 0), ez - 1);
  const long long w = (((long long)c * ex + ii) * ey + jj) * ez + kk;
  const double e = eps[w];
  const double s = mask[((long long)c * g.X + i) * g.Y + j];
  const double cb = dt / (e + 0.5 * s * dt);
  atomicAdd(&grad[w], (float)(cb / dt * (double)G[t]));
}

hipError_t launch_grad_eps(const Geom& g, const pjz_fdtd_problem& p, const float* eps,
                           const float* mask, const char* ws, const Layout& L, float* grad_eps,
                           hipStream_t s) {
  hipError_t e = launch_zero(grad_eps, 3 * (size_t)p.ex * p.ey * p.ez, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_grad_eps, dim3(nblocks(3 * g.N, 256)), dim3(256), 0, s, eps, mask,
                     (const float*)(ws + L.accG), grad_eps, g, p.x0, p.y0, p.z0, p.ex, p.ey,
                     p.ez, (double)p.dt);
  return hipGetLastError();
}

}  // namespace pjz
