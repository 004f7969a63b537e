// pjz's density() (src/pjz/_density.py:48-79) on the device: the three-field
// scheme's cone blur of the design variable (edge-padded, `_filter`, :17-20),
// the projection that binarises every pixel off the eta boundary
// (`_project`, :23-31) and the minimum-feature-size loss weighted by the
// inflection indicator exp(-c |central difference|^2) (:34-45), with the
// exact reverse pass for dL/du.
//
// Layout: nb variables of (X, Y), row-major (y fastest), fp32.
//   forward:  k_density_conv4<false>  uf = cone * pad_edge(u) (LDS tile +
//                              halo, four outputs a lane; k_density_blur,
//                              one output a lane, above half-width 24)
//             k_density_finish dens, loss from uf         (one pixel a lane)
//   backward: k_density_gradf  guf = dL/duf               (pixel-local terms
//                              plus the inflection's neighbour terms)
//             k_density_conv4<true> / k_density_corr  gpad = K * guf on
//                              the padded domain
//             k_density_fold   gu = gpad folded back through the edge pad
// The blur is (2r+1)^2 multiply-adds a pixel against 8 bytes of HBM
// traffic, so at the radii pjz uses it is VALU-bound, not HBM-bound.
#include <cmath>

#include "fdtd_common.h"

namespace pjz {

namespace {

constexpr int kDenTile = 16;  // 16 x 16 pixels a workgroup (256 lanes)

// Cone weight before normalisation: max(0, radius - |(a, b)|) (:9-14).
__device__ inline float cone_weight(int a, int b, float radius) {
  return fmaxf(0.0f, radius - sqrtf((float)(a * a + b * b)));
}

__global__ __launch_bounds__(256) void k_density_blur(const float* __restrict__ u,
                                                      float* __restrict__ uf, int X, int Y,
                                                      int r, float radius, float inv_sum) {
  extern __shared__ float sm[];
  const int S = kDenTile + 2 * r, K = 2 * r + 1;
  float* su = sm;          // S x S tile of u with an r-pixel edge-clamped halo
  float* sw = sm + S * S;  // K x K normalised cone
  const size_t b = (size_t)blockIdx.z * X * Y;
  const int x0 = blockIdx.x * kDenTile, y0 = blockIdx.y * kDenTile;
  for (int t = threadIdx.x; t < S * S; t += blockDim.x) {
    const int xi = min(max(x0 + t / S - r, 0), X - 1);
    const int yi = min(max(y0 + t % S - r, 0), Y - 1);
    su[t] = u[b + (size_t)xi * Y + yi];
  }
  for (int t = threadIdx.x; t < K * K; t += blockDim.x)
    sw[t] = cone_weight(t / K - r, t % K - r, radius) * inv_sum;
  __syncthreads();
  const int tx = threadIdx.x / kDenTile, ty = threadIdx.x % kDenTile;
  const int x = x0 + tx, y = y0 + ty;
  if (x >= X || y >= Y) return;
  float acc = 0.0f;
  for (int a = 0; a < K; ++a) {
    const float* row = su + (tx + a) * S + ty;
    const float* w = sw + a * K;
    for (int c = 0; c < K; ++c) acc = fmaf(w[c], row[c], acc);
  }
  uf[b + (size_t)x * Y + y] = acc;
}

// The same blur with four outputs a lane along y (16 x 64 outputs a
// workgroup): a lane slides a K + 3 window along each LDS row of the tile, so
// each value it reads serves four outputs (~3.4 x fewer LDS reads at K = 9).
// Cone half-widths up to kDenR4 (the tile and its halo within 64 KB of LDS).
// Two uses: the forward blur (kPad false: output (x, y) of X x Y reads the
// edge-clamped input at x - r .. x + r) and the reverse pass's correlation
// over the padded domain (kPad true: output (p, q) of (X + 2r) x (Y + 2r)
// reads the input at p - 2r .. p, zero outside it; the cone is symmetric, so
// the same weights apply).
constexpr int kDenTileY4 = 64;
constexpr int kDenR4 = 24;

template <bool kPad>
__global__ __launch_bounds__(256) void k_density_conv4(const float* __restrict__ u,
                                                       float* __restrict__ out, int X, int Y,
                                                       int r, float radius, float inv_sum) {
  extern __shared__ float sm[];
  const int SX = kDenTile + 2 * r, SY = kDenTileY4 + 2 * r, K = 2 * r + 1;
  const int OX = kPad ? X + 2 * r : X, OY = kPad ? Y + 2 * r : Y;
  const int off = kPad ? 2 * r : r;
  float* su = sm;            // SX x SY input tile with its halo
  float* sw = sm + SX * SY;  // K x K normalised cone
  const float* in = u + (size_t)blockIdx.z * X * Y;
  const int x0 = blockIdx.x * kDenTile, y0 = blockIdx.y * kDenTileY4;
  for (int t = threadIdx.x; t < SX * SY; t += blockDim.x) {
    const int xi = x0 + t / SY - off, yi = y0 + t % SY - off;
    if (kPad)
      su[t] = xi >= 0 && xi < X && yi >= 0 && yi < Y ? in[(size_t)xi * Y + yi] : 0.0f;
    else
      su[t] = in[(size_t)min(max(xi, 0), X - 1) * Y + min(max(yi, 0), Y - 1)];
  }
  for (int t = threadIdx.x; t < K * K; t += blockDim.x)
    sw[t] = cone_weight(t / K - r, t % K - r, radius) * inv_sum;
  __syncthreads();
  const int tx = threadIdx.x / 16, ty = (threadIdx.x % 16) * 4;
  const int x = x0 + tx, y = y0 + ty;
  if (x >= OX || y >= OY) return;
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
  for (int a = 0; a < K; ++a) {
    const float* row = su + (tx + a) * SY + ty;
    const float* w = sw + a * K;
    float v0 = row[0], v1 = row[1], v2 = row[2];
    for (int c = 0; c < K; ++c) {
      const float v3 = row[c + 3];
      const float wc = w[c];
      a0 = fmaf(wc, v0, a0);
      a1 = fmaf(wc, v1, a1);
      a2 = fmaf(wc, v2, a2);
      a3 = fmaf(wc, v3, a3);
      v0 = v1;
      v1 = v2;
      v2 = v3;
    }
  }
  float* o = out + (size_t)blockIdx.z * OX * OY + (size_t)x * OY + y;
  o[0] = a0;
  if (y + 1 < OY) o[1] = a1;
  if (y + 2 < OY) o[2] = a2;
  if (y + 3 < OY) o[3] = a3;
}

// Launches k_density_conv4 over an OX x OY output.
template <bool kPad>
void launch_conv4(const float* in, float* out, int nb, int X, int Y, int r, float radius,
                  float inv, hipStream_t s) {
  const int OX = kPad ? X + 2 * r : X, OY = kPad ? Y + 2 * r : Y;
  const size_t lds =
      sizeof(float) * (size_t)((kDenTile + 2 * r) * (kDenTileY4 + 2 * r) + (2 * r + 1) * (2 * r + 1));
  const dim3 grid((OX + kDenTile - 1) / kDenTile, (OY + kDenTileY4 - 1) / kDenTileY4, nb);
  hipLaunchKernelGGL(k_density_conv4<kPad>, grid, dim3(256), lds, s, in, out, X, Y, r, radius, inv);
}

// Per-pixel quantities of the projection and loss at (x, y) of one variable.
struct DenPix {
  float f, proj, w, A, B, dx, dy;  // uf, projected, inflection weight, min(0, f - hi),
                                   // min(0, lo - f), central differences
  bool bnd;                        // on the eta boundary
};

__device__ inline DenPix den_pixel(const float* __restrict__ uf, int X, int Y, int x, int y,
                                   float c, float eta, float lo, float hi) {
  DenPix p;
  const int xm = max(x - 1, 0), xp = min(x + 1, X - 1);
  const int ym = max(y - 1, 0), yp = min(y + 1, Y - 1);
  p.f = uf[(size_t)x * Y + y];
  const float fxm = uf[(size_t)xm * Y + y], fxp = uf[(size_t)xp * Y + y];
  const float fym = uf[(size_t)x * Y + ym], fyp = uf[(size_t)x * Y + yp];
  // Boundary: the sign of uf - eta >= 0 differs from a 4-neighbour's (edges
  // replicated, so an edge pixel's missing neighbour is itself) (:23-27).
  const bool s = p.f - eta >= 0.0f;
  p.bnd = (s != (fxm - eta >= 0.0f)) || (s != (fxp - eta >= 0.0f)) ||
          (s != (fym - eta >= 0.0f)) || (s != (fyp - eta >= 0.0f));
  const float d = p.f - eta;
  const float sgn = d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f);
  p.proj = p.bnd ? p.f : (sgn + 1.0f) * 0.5f;  // (:30-31)
  p.dx = fxp - fxm;
  p.dy = fyp - fym;
  p.w = expf(-c * (p.dx * p.dx + p.dy * p.dy));  // (:34-38)
  p.A = fminf(0.0f, p.f - hi);
  p.B = fminf(0.0f, lo - p.f);
  return p;
}

__global__ __launch_bounds__(256) void k_density_finish(const float* __restrict__ uf,
                                                        float* __restrict__ dens,
                                                        float* __restrict__ loss, int X, int Y,
                                                        float alpha, float c, float eta, float lo,
                                                        float hi) {
  const int x = blockIdx.x * kDenTile + threadIdx.x / kDenTile;
  const int y = blockIdx.y * kDenTile + threadIdx.x % kDenTile;
  if (x >= X || y >= Y) return;
  const size_t b = (size_t)blockIdx.z * X * Y;
  const DenPix p = den_pixel(uf + b, X, Y, x, y, c, eta, lo, hi);
  const size_t i = b + (size_t)x * Y + y;
  dens[i] = alpha * p.proj + (1.0f - alpha) * p.f;
  // (:41-45)
  loss[i] = p.proj * p.w * (p.A * p.A) + (1.0f - p.proj) * p.w * (p.B * p.B);
}

// dL/duf.  Pixel-local: dens = alpha proj + (1 - alpha) f with dproj/df = 1
// on the boundary and 0 off it (the sign's derivative); loss = w (proj A^2 +
// (1 - proj) B^2) with dA^2/df = 2A, dB^2/df = -2B.  Through the inflection
// weight w_q = exp(-c (dx_q^2 + dy_q^2)) of the pixels q whose central
// differences read this pixel: dL/dw_q dw_q/duf = -2 c w_q gl_q (proj_q A_q^2
// + (1 - proj_q) B_q^2) (dx_q or dy_q) (+1 where this pixel is q's clamped
// upper neighbour, -1 where it is the lower one; both at a clamped edge).
__global__ __launch_bounds__(256) void k_density_gradf(
    const float* __restrict__ uf, const float* __restrict__ gd, const float* __restrict__ gl,
    float* __restrict__ guf, int X, int Y, float alpha, float c, float eta, float lo, float hi) {
  const int x = blockIdx.x * kDenTile + threadIdx.x / kDenTile;
  const int y = blockIdx.y * kDenTile + threadIdx.x % kDenTile;
  if (x >= X || y >= Y) return;
  const size_t b = (size_t)blockIdx.z * X * Y;
  const float* f = uf + b;
  const size_t i = b + (size_t)x * Y + y;
  const DenPix p = den_pixel(f, X, Y, x, y, c, eta, lo, hi);
  const float dp = p.bnd ? 1.0f : 0.0f;
  float g = gd[i] * (alpha * dp + (1.0f - alpha));
  g += gl[i] * p.w * (dp * (p.A * p.A - p.B * p.B) + 2.0f * p.proj * p.A -
                      2.0f * (1.0f - p.proj) * p.B);
  // Neighbour terms along x: pixels q = (qx, y), qx in x-1 .. x+1.
  for (int qx = max(x - 1, 0); qx <= min(x + 1, X - 1); ++qx) {
    const int up = min(qx + 1, X - 1), dn = max(qx - 1, 0);
    const float sgn = (up == x ? 1.0f : 0.0f) - (dn == x ? 1.0f : 0.0f);
    if (sgn == 0.0f) continue;
    const DenPix q = den_pixel(f, X, Y, qx, y, c, eta, lo, hi);
    const float dLdw = gl[b + (size_t)qx * Y + y] *
                       (q.proj * (q.A * q.A) + (1.0f - q.proj) * (q.B * q.B));
    g += -2.0f * c * q.w * dLdw * q.dx * sgn;
  }
  for (int qy = max(y - 1, 0); qy <= min(y + 1, Y - 1); ++qy) {
    const int up = min(qy + 1, Y - 1), dn = max(qy - 1, 0);
    const float sgn = (up == y ? 1.0f : 0.0f) - (dn == y ? 1.0f : 0.0f);
    if (sgn == 0.0f) continue;
    const DenPix q = den_pixel(f, X, Y, x, qy, c, eta, lo, hi);
    const float dLdw = gl[b + (size_t)x * Y + qy] *
                       (q.proj * (q.A * q.A) + (1.0f - q.proj) * (q.B * q.B));
    g += -2.0f * c * q.w * dLdw * q.dy * sgn;
  }
  guf[i] = g;
}

// gpad(p, q) = sum over in-range (i, j) of K(p - i, q - j) guf(i, j), on the
// padded domain p in [-r, X + r), q in [-r, Y + r) (stored at p + r, q + r).
__global__ __launch_bounds__(256) void k_density_corr(const float* __restrict__ guf,
                                                      float* __restrict__ gpad, int X, int Y,
                                                      int r, float radius, float inv_sum) {
  extern __shared__ float sm[];
  const int K = 2 * r + 1;
  for (int t = threadIdx.x; t < K * K; t += blockDim.x)
    sm[t] = cone_weight(t / K - r, t % K - r, radius) * inv_sum;
  __syncthreads();
  const int PX = X + 2 * r, PY = Y + 2 * r;
  const int pp = blockIdx.x * kDenTile + threadIdx.x / kDenTile;
  const int qq = blockIdx.y * kDenTile + threadIdx.x % kDenTile;
  if (pp >= PX || qq >= PY) return;
  const int p = pp - r, q = qq - r;
  const float* g = guf + (size_t)blockIdx.z * X * Y;
  float acc = 0.0f;
  for (int i = max(p - r, 0); i <= min(p + r, X - 1); ++i) {
    const float* w = sm + (p - i + r) * K;
    for (int j = max(q - r, 0); j <= min(q + r, Y - 1); ++j)
      acc = fmaf(w[q - j + r], g[(size_t)i * Y + j], acc);
  }
  gpad[(size_t)blockIdx.z * PX * PY + (size_t)pp * PY + qq] = acc;
}

// gu(x, y) = sum of gpad over the padded positions the edge pad maps to
// (x, y): itself inside, the r pads beyond it on an edge, (r+1)^2 at a corner.
__global__ __launch_bounds__(256) void k_density_fold(const float* __restrict__ gpad,
                                                      float* __restrict__ gu, int X, int Y,
                                                      int r) {
  const int x = blockIdx.x * kDenTile + threadIdx.x / kDenTile;
  const int y = blockIdx.y * kDenTile + threadIdx.x % kDenTile;
  if (x >= X || y >= Y) return;
  const int PX = X + 2 * r, PY = Y + 2 * r;
  const int p0 = x == 0 ? 0 : x + r, p1 = x == X - 1 ? PX - 1 : x + r;
  const int q0 = y == 0 ? 0 : y + r, q1 = y == Y - 1 ? PY - 1 : y + r;
  const float* g = gpad + (size_t)blockIdx.z * PX * PY;
  float acc = 0.0f;
  for (int p = p0; p <= p1; ++p)
    for (int q = q0; q <= q1; ++q) acc += g[(size_t)p * PY + q];
  gu[(size_t)blockIdx.z * X * Y + (size_t)x * Y + y] = acc;
}

dim3 den_grid(int X, int Y, int nb) {
  return dim3((X + kDenTile - 1) / kDenTile, (Y + kDenTile - 1) / kDenTile, nb);
}

}  // namespace

int density_half_width(float radius) { return (int)std::ceil((double)radius - 0.5); }

// 1 / sum of the cone's weights, summed in double (the normalisation of :14).
static float cone_inv_sum(float radius, int r) {
  double s = 0.0;
  for (int a = -r; a <= r; ++a)
    for (int b = -r; b <= r; ++b) {
      const double w = (double)radius - std::sqrt((double)(a * a + b * b));
      if (w > 0.0) s += w;
    }
  return (float)(1.0 / s);
}

size_t density_workspace_bytes(int nb, int X, int Y, float radius) {
  const int r = density_half_width(radius);
  return align256(sizeof(float) * (size_t)nb * X * Y) +
         align256(sizeof(float) * (size_t)nb * (X + 2 * r) * (Y + 2 * r));
}

hipError_t density_forward(int nb, int X, int Y, float radius, float alpha, float c, float eta,
                           float lo, float hi, const float* u, float* dens, float* loss,
                           float* uf, hipStream_t s) {
  const int r = density_half_width(radius);
  const float inv = cone_inv_sum(radius, r);
  const int S = kDenTile + 2 * r, K = 2 * r + 1;
  if (r <= kDenR4) {
    launch_conv4<false>(u, uf, nb, X, Y, r, radius, inv, s);
  } else {
    const size_t lds = sizeof(float) * (size_t)(S * S + K * K);
    hipLaunchKernelGGL(k_density_blur, den_grid(X, Y, nb), dim3(256), lds, s, u, uf, X, Y, r,
                       radius, inv);
  }
  hipLaunchKernelGGL(k_density_finish, den_grid(X, Y, nb), dim3(256), 0, s, uf, dens, loss, X, Y,
                     alpha, c, eta, lo, hi);
  return hipGetLastError();
}

hipError_t density_backward(int nb, int X, int Y, float radius, float alpha, float c, float eta,
                            float lo, float hi, const float* uf, const float* gd,
                            const float* gl, float* gu, char* ws, hipStream_t s) {
  const int r = density_half_width(radius);
  const float inv = cone_inv_sum(radius, r);
  float* guf = (float*)ws;
  float* gpad = (float*)(ws + align256(sizeof(float) * (size_t)nb * X * Y));
  const int K = 2 * r + 1;
  hipLaunchKernelGGL(k_density_gradf, den_grid(X, Y, nb), dim3(256), 0, s, uf, gd, gl, guf, X, Y,
                     alpha, c, eta, lo, hi);
  if (r <= kDenR4)
    launch_conv4<true>(guf, gpad, nb, X, Y, r, radius, inv, s);
  else
    hipLaunchKernelGGL(k_density_corr, den_grid(X + 2 * r, Y + 2 * r, nb), dim3(256),
                       sizeof(float) * (size_t)(K * K), s, guf, gpad, X, Y, r, radius, inv);
  hipLaunchKernelGGL(k_density_fold, den_grid(X, Y, nb), dim3(256), 0, s, gpad, gu, X, Y, r);
  return hipGetLastError();
}

}  // namespace pjz
