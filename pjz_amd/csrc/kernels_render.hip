// pjz's permittivity render (src/pjz/_epsilon.py:10-108) on the device:
// layered 2-D profiles magnified 2m x 2m -> (3, xx, yy, zz) Yee-cell eps by
// Farjadpour et al.'s diagonal sub-pixel average, with the exact reverse pass
// for the layers and the interface positions.
//
//   forward:  k_render_stats    per (component, layer, pixel): mean, mean of
//                               the inverse and the two in-plane moments of
//                               the component's half-cell-shifted 2m x 2m
//                               tile (:12-16, 26-45)
//             k_render_shares   per (z grid, layer, z): the layer's share of
//                               the cell and its z moment (:47-61, 75-87)
//             k_render_combine  per (component, x, y, z): the moments mixed
//                               by the shares into 1 / (p aoi + (1 - p) /
//                               avg) or the plain average (:63-99)
//   backward: k_render_bwd_stats   dL/dstats, summed over z
//             k_render_bwd_share   dL/d(z share, z-moment share), summed over
//                                  chunks of x, y (workgroups per component,
//                                  z and chunk)
//             k_render_bwd_pos     dL/d(interface positions)
//             k_render_bwd_layers  dL/d(layer sub-pixels), gathered through
//                                  the shifts' edge replication
// Layouts (fp32): layers (ll, SX = 2m xx, SY = 2m yy); stats (3, 4, ll, xx,
// yy) with the four moments (mean, mean of inverse, x moment, y moment),
// then each layer's share of each cell (k_render_shares); eps (3, xx, yy,
// zz).  Grid bounds gs, ge are (zz, 2): column 0 for the x
// and y components, column 1 for z (:18-24).
#include <algorithm>
#include <cmath>

#include "fdtd_common.h"

namespace pjz {

namespace {

struct RenderDims {
  int ll, m, xx, yy, zz, simple;
};

__device__ inline size_t stat_at(const RenderDims& d, int a, int q, int l, int x, int y) {
  return ((((size_t)a * 4 + q) * d.ll + l) * d.xx + x) * (size_t)d.yy + y;
}

// Source sub-pixel a component's shifted sub-pixel (i, j) reads: components
// not along x are shifted by m sub-pixels in x (edge-replicated at the
// start), likewise in y (:12-16).
__device__ inline int src_x(int a, int i, int m) { return a != 0 ? max(i - m, 0) : i; }
__device__ inline int src_y(int a, int j, int m) { return a != 1 ? max(j - m, 0) : j; }

__global__ __launch_bounds__(256) void k_render_stats(RenderDims d, const float* __restrict__ L,
                                                      float* __restrict__ st) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long n_out = 3LL * d.ll * d.xx * d.yy;
  if (i >= n_out) return;
  const int y = (int)(i % d.yy), x = (int)((i / d.yy) % d.xx);
  const int l = (int)((i / ((long long)d.yy * d.xx)) % d.ll), a = (int)(i / ((long long)d.yy * d.xx * d.ll));
  const int n = 2 * d.m, SX = n * d.xx, SY = n * d.yy;
  const float* Ll = L + (size_t)l * SX * SY;
  float s = 0.0f, si = 0.0f, sx = 0.0f, sy = 0.0f;
  for (int di = 0; di < n; ++di) {
    const float wx = (di - (d.m - 0.5f)) / (float)(n * n);
    const int r = src_x(a, n * x + di, d.m);
    for (int dj = 0; dj < n; ++dj) {
      const float wy = (dj - (d.m - 0.5f)) / (float)(n * n);
      const float v = Ll[(size_t)r * SY + src_y(a, n * y + dj, d.m)];
      s += v;
      si += 1.0f / v;
      sx += v * wx;
      sy += v * wy;
    }
  }
  const float inv = 1.0f / (float)(n * n);
  st[stat_at(d, a, 0, l, x, y)] = s * inv;
  st[stat_at(d, a, 1, l, x, y)] = si * inv;
  // mean of 12 (2m) v w (:35-37)
  st[stat_at(d, a, 2, l, x, y)] = 12.0f * n * sx * inv;
  st[stat_at(d, a, 3, l, x, y)] = 12.0f * n * sy * inv;
}

// Layer l's share u of cell z (grid column col) and its z moment
// uz = u zc / (dz^2 / 12) (:47-61, 75-87), with what the reverse pass needs.
struct Share {
  float u, uz, p0, p1, gs, ge;
};

__device__ inline Share layer_share(const RenderDims& d, const float* __restrict__ pos,
                                    const float* __restrict__ gsa, const float* __restrict__ gea,
                                    int l, int z, int col) {
  Share s;
  s.gs = gsa[2 * z + col];
  s.ge = gea[2 * z + col];
  const float lo = l == 0 ? -INFINITY : pos[l - 1];
  const float hi = l == d.ll - 1 ? INFINITY : pos[l];
  s.p0 = fminf(fmaxf(lo, s.gs), s.ge);
  s.p1 = fminf(fmaxf(hi, s.gs), s.ge);
  const float dz = s.ge - s.gs;
  s.u = (s.p1 - s.p0) / dz;
  const float zc = (s.p0 + s.p1) / 2 - (s.gs + s.ge) / 2;
  s.uz = s.u * zc / (dz * dz / 12.0f);
  return s;
}

// The share table: (u, uz) of every layer and cell for both z grids,
// table[((col * ll + l) * zz + z) * 2 + {0, 1}], after the 12 ll xx yy
// moments in the stats array.
__global__ __launch_bounds__(256) void k_render_shares(RenderDims d, const float* __restrict__ pos,
                                                       const float* __restrict__ gsa,
                                                       const float* __restrict__ gea,
                                                       float* __restrict__ tab) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * d.ll * d.zz) return;
  const int z = i % d.zz, l = (i / d.zz) % d.ll, col = i / (d.zz * d.ll);
  const Share sh = layer_share(d, pos, gsa, gea, l, z, col);
  tab[2 * i] = sh.u;
  tab[2 * i + 1] = sh.uz;
}

__device__ inline const float* share_at(const RenderDims& d, const float* tab, int a, int l,
                                        int z) {
  return tab + (((a == 2 ? 1 : 0) * d.ll + l) * (size_t)d.zz + z) * 2;
}

// Mixed sums of one output cell and its value.
struct Mix {
  float A, B, G[3], out, pii, T;
};

__device__ inline Mix mix_cell(const RenderDims& d, const float* __restrict__ st,
                               const float* __restrict__ tab, int a, int x, int y, int z) {
  Mix r;
  r.A = r.B = r.G[0] = r.G[1] = r.G[2] = 0.0f;
  for (int l = 0; l < d.ll; ++l) {
    const float* sh = share_at(d, tab, a, l, z);
    const float u = sh[0], uz = sh[1];
    const float avg = st[stat_at(d, a, 0, l, x, y)];
    r.A += avg * u;
    r.B += st[stat_at(d, a, 1, l, x, y)] * u;
    r.G[0] += st[stat_at(d, a, 2, l, x, y)] * u;
    r.G[1] += st[stat_at(d, a, 3, l, x, y)] * u;
    r.G[2] += avg * uz;
  }
  if (d.simple) {
    r.out = r.A;
    r.pii = r.T = 0.0f;
    return r;
  }
  r.T = r.G[0] * r.G[0] + r.G[1] * r.G[1] + r.G[2] * r.G[2];
  // (the component's moment selected, not indexed: G stays in registers)
  const float ga = a == 0 ? r.G[0] : (a == 1 ? r.G[1] : r.G[2]);
  r.pii = ga * ga / (r.T == 0.0f ? 1.0f : r.T);  // (:92-95)
  r.out = 1.0f / (r.pii * r.B + (1.0f - r.pii) * (1.0f / r.A));  // (:97)
  return r;
}

__global__ __launch_bounds__(256) void k_render_combine(RenderDims d, const float* __restrict__ st,
                                                        const float* __restrict__ tab,
                                                        float* __restrict__ eps) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= 3LL * d.xx * d.yy * d.zz) return;
  const int z = (int)(i % d.zz), y = (int)((i / d.zz) % d.yy), x = (int)((i / ((long long)d.zz * d.yy)) % d.xx);
  const int a = (int)(i / ((long long)d.zz * d.yy * d.xx));
  eps[i] = mix_cell(d, st, tab, a, x, y, z).out;
}

// Cotangents of a cell's mixed sums A, B, G[0..2] from g = dL/d(out).
struct MixGrad {
  float gA, gB, gG[3];
};

__device__ inline MixGrad mix_grad(const RenderDims& d, const Mix& r, int a, float g) {
  MixGrad t;
  if (d.simple) {
    t.gA = g;
    t.gB = t.gG[0] = t.gG[1] = t.gG[2] = 0.0f;
    return t;
  }
  // out = 1 / D, D = pii B + (1 - pii) / A, pii = G_a^2 / T (T = sum G_k^2,
  // no gradient through T where T = 0, where every G_k and pii are 0).
  const float gD = -g * r.out * r.out;
  const float iA = 1.0f / r.A;
  t.gA = gD * (1.0f - r.pii) * (-iA * iA);
  t.gB = gD * r.pii;
  const float gp = gD * (r.B - iA);
  const float ga = a == 0 ? r.G[0] : (a == 1 ? r.G[1] : r.G[2]);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float dp = 0.0f;
    if (r.T != 0.0f) {
      dp = -ga * ga * 2.0f * r.G[k] / (r.T * r.T);
      if (k == a) dp += 2.0f * ga / r.T;
    }
    t.gG[k] = gp * dp;
  }
  return t;
}

// dL/dstats (3, 4, ll, xx, yy): sum over z of the cell cotangents times the
// shares.  One thread per (component, x, y), up to kRenderMaxLayers layers.
constexpr int kRenderMaxLayers = 16;
static_assert(kRenderMaxLayers == PJZ_FDTD_RENDER_MAX_LAYERS, "layer bound of the C ABI");

__global__ __launch_bounds__(256) void k_render_bwd_stats(RenderDims d, const float* __restrict__ st,
                                                          const float* __restrict__ tab,
                                                          const float* __restrict__ geps,
                                                          float* __restrict__ gst) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= 3LL * d.xx * d.yy) return;
  const int y = (int)(i % d.yy), x = (int)((i / d.yy) % d.xx), a = (int)(i / ((long long)d.yy * d.xx));
  float acc[kRenderMaxLayers][4];
  for (int l = 0; l < kRenderMaxLayers; ++l) acc[l][0] = acc[l][1] = acc[l][2] = acc[l][3] = 0.0f;
  for (int z = 0; z < d.zz; ++z) {
    const Mix r = mix_cell(d, st, tab, a, x, y, z);
    const float g = geps[(((size_t)a * d.xx + x) * d.yy + y) * d.zz + z];
    const MixGrad t = mix_grad(d, r, a, g);
#pragma unroll
    for (int l = 0; l < kRenderMaxLayers; ++l) {
      if (l < d.ll) {
        const float* sh = share_at(d, tab, a, l, z);
        acc[l][0] += t.gA * sh[0] + t.gG[2] * sh[1];
        acc[l][1] += t.gB * sh[0];
        acc[l][2] += t.gG[0] * sh[0];
        acc[l][3] += t.gG[1] * sh[0];
      }
    }
  }
#pragma unroll
  for (int l = 0; l < kRenderMaxLayers; ++l) {
    if (l < d.ll) {
#pragma unroll
      for (int q = 0; q < 4; ++q) gst[stat_at(d, a, q, l, x, y)] = acc[l][q];
    }
  }
}

// dL/du and dL/duz (3, ll, zz), in parts: workgroup (z, a, ch) sums chunk ch
// of the cells of its xy plane into part[((ch * 3 + a) * ll + l) * zz + z]
// (u) and the same after nch * 3 * ll * zz floats (uz); k_render_sum_parts
// adds the chunks.
__device__ inline int render_chunks(const RenderDims& d) {
  return min(max((d.xx * d.yy + 4095) / 4096, 1), 64);
}

__global__ __launch_bounds__(256) void k_render_bwd_share(RenderDims d, const float* __restrict__ st,
                                                          const float* __restrict__ tab,
                                                          const float* __restrict__ geps,
                                                          float* __restrict__ part) {
  __shared__ float red[256];
  const int z = blockIdx.x, a = blockIdx.y, ch = blockIdx.z, nch = gridDim.z;
  const int np = d.xx * d.yy;
  const int p0 = (int)((long long)ch * np / nch), p1 = (int)((long long)(ch + 1) * np / nch);
  float au[kRenderMaxLayers], auz[kRenderMaxLayers];
  for (int l = 0; l < kRenderMaxLayers; ++l) au[l] = auz[l] = 0.0f;
  for (int p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    const int x = p / d.yy, y = p % d.yy;
    const Mix r = mix_cell(d, st, tab, a, x, y, z);
    const MixGrad t = mix_grad(d, r, a, geps[(((size_t)a * d.xx + x) * d.yy + y) * d.zz + z]);
#pragma unroll
    for (int l = 0; l < kRenderMaxLayers; ++l) {
      if (l < d.ll) {
        const float avg = st[stat_at(d, a, 0, l, x, y)];
        au[l] += t.gA * avg + t.gB * st[stat_at(d, a, 1, l, x, y)] +
                 t.gG[0] * st[stat_at(d, a, 2, l, x, y)] +
                 t.gG[1] * st[stat_at(d, a, 3, l, x, y)];
        auz[l] += t.gG[2] * avg;
      }
    }
  }
#pragma unroll
  for (int l = 0; l < kRenderMaxLayers; ++l) {
    if (l < d.ll) {
#pragma unroll
      for (int w = 0; w < 2; ++w) {
        red[threadIdx.x] = w == 0 ? au[l] : auz[l];
        __syncthreads();
        for (int h = blockDim.x / 2; h > 0; h >>= 1) {
          if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
          __syncthreads();
        }
        if (threadIdx.x == 0)
          part[(size_t)w * nch * 3 * d.ll * d.zz + (((size_t)ch * 3 + a) * d.ll + l) * d.zz + z] =
              red[0];
        __syncthreads();
      }
    }
  }
}

// d clip(v, lo, hi) / dv with torch's (and jax's) halves at ties.
__device__ inline float dclip(float v, float lo, float hi) {
  const float m = fmaxf(v, lo);
  const float a = v > lo ? 1.0f : (v == lo ? 0.5f : 0.0f);
  const float b = m < hi ? 1.0f : (m == hi ? 0.5f : 0.0f);
  return a * b;
}

// dL/dpos (ll - 1): interface j is layer j's upper bound and layer j+1's
// lower bound.  One workgroup; lane j of the first ll - 1 sums over
// components and z.
// Adds the chunks of k_render_bwd_share: gus[w][a][l][z] (w = 0: dL/du,
// 1: dL/duz), one lane per entry.
__global__ __launch_bounds__(256) void k_render_sum_parts(RenderDims d, const float* __restrict__ part,
                                                          float* __restrict__ gus) {
  const int n = 3 * d.ll * d.zz;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * n) return;
  const int w = i / n, e = i % n;
  const int nch = render_chunks(d);
  float t = 0.0f;
  for (int ch = 0; ch < nch; ++ch) t += part[((size_t)w * nch + ch) * n + e];
  gus[i] = t;
}

__global__ __launch_bounds__(64) void k_render_bwd_pos(RenderDims d, const float* __restrict__ pos,
                                                       const float* __restrict__ gsa,
                                                       const float* __restrict__ gea,
                                                       const float* __restrict__ gus,
                                                       float* __restrict__ gpos) {
  const size_t n = (size_t)3 * d.ll * d.zz;
  // dL/du (w = 0) and dL/duz (w = 1) of layer l at (a, z).
  auto sum = [&](int w, int a, int l, int z) {
    return gus[w * n + ((size_t)a * d.ll + l) * d.zz + z];
  };
  for (int j = threadIdx.x; j < d.ll - 1; j += blockDim.x) {
    float acc = 0.0f;
    for (int a = 0; a < 3; ++a) {
      const int col = a == 2 ? 1 : 0;
      for (int z = 0; z < d.zz; ++z) {
        const float gs = gsa[2 * z + col], ge = gea[2 * z + col];
        const float dz = ge - gs, c = (gs + ge) / 2, k = 12.0f / (dz * dz * dz);
        // layer j: hi = pos[j] -> p1
        {
          const Share sh = layer_share(d, pos, gsa, gea, j, z, col);
          const float dp1 = dclip(pos[j], gs, ge);
          acc += (sum(0, a, j, z) / dz + sum(1, a, j, z) * k * (sh.p1 - c)) * dp1;
        }
        // layer j + 1: lo = pos[j] -> p0
        {
          const Share sh = layer_share(d, pos, gsa, gea, j + 1, z, col);
          const float dp0 = dclip(pos[j], gs, ge);
          acc += (-sum(0, a, j + 1, z) / dz + sum(1, a, j + 1, z) * k * (c - sh.p0)) * dp0;
        }
      }
    }
    gpos[j] = acc;
  }
}

// dL/dlayers (ll, SX, SY): each source sub-pixel gathers, for every
// component, the shifted sub-pixels that read it (its own position, m on
// from it, or the first m + 1 along a shifted edge; none for the last m of a
// shifted axis) with their tile's moment cotangents: d mean = 1 / n^2,
// d mean of inverse = -1 / (v^2 n^2), d moments = 12 n w / n^2.
__global__ __launch_bounds__(256) void k_render_bwd_layers(RenderDims d, const float* __restrict__ L,
                                                           const float* __restrict__ gst,
                                                           float* __restrict__ gL) {
  const int n = 2 * d.m, SX = n * d.xx, SY = n * d.yy;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= (long long)d.ll * SX * SY) return;
  const int sj = (int)(i % SY), si = (int)((i / SY) % SX), l = (int)(i / ((long long)SY * SX));
  const float v = L[i];
  const float inv = 1.0f / (float)(n * n);
  float acc = 0.0f;
  for (int a = 0; a < 3; ++a) {
    // shifted positions along x reading source row si
    int i0, i1;
    if (a == 0) {
      i0 = i1 = si;
    } else if (si == 0) {
      i0 = 0;
      i1 = min(d.m, SX - 1);
    } else {
      i0 = i1 = si + d.m;
      if (i0 >= SX) continue;
    }
    int j0, j1;
    if (a == 1) {
      j0 = j1 = sj;
    } else if (sj == 0) {
      j0 = 0;
      j1 = min(d.m, SY - 1);
    } else {
      j0 = j1 = sj + d.m;
      if (j0 >= SY) continue;
    }
    for (int ii = i0; ii <= i1; ++ii) {
      const int x = ii / n, di = ii % n;
      const float wx = (di - (d.m - 0.5f)) / (float)(n * n);
      for (int jj = j0; jj <= j1; ++jj) {
        const int y = jj / n, dj = jj % n;
        const float wy = (dj - (d.m - 0.5f)) / (float)(n * n);
        acc += (gst[stat_at(d, a, 0, l, x, y)] - gst[stat_at(d, a, 1, l, x, y)] / (v * v) +
                gst[stat_at(d, a, 2, l, x, y)] * 12.0f * n * wx +
                gst[stat_at(d, a, 3, l, x, y)] * 12.0f * n * wy) * inv;
      }
    }
  }
  gL[i] = acc;
}

unsigned blocks_for(long long n, int threads = 256) { return (unsigned)((n + threads - 1) / threads); }

}  // namespace

int render_max_layers() { return kRenderMaxLayers; }

size_t render_stats_floats(int ll, int xx, int yy, int zz) {
  return (size_t)12 * ll * xx * yy + (size_t)4 * ll * zz;
}

static int render_chunks_host(int xx, int yy) {
  return std::min(std::max((xx * yy + 4095) / 4096, 1), 64);
}

size_t render_workspace_bytes(int ll, int xx, int yy, int zz) {
  return align256(sizeof(float) * (size_t)12 * ll * xx * yy) +
         align256(sizeof(float) * (size_t)2 * render_chunks_host(xx, yy) * 3 * ll * zz) +
         align256(sizeof(float) * (size_t)2 * 3 * ll * zz);
}

hipError_t render_forward(int ll, int m, int xx, int yy, int zz, int simple, const float* layers,
                          const float* pos, const float* gs, const float* ge, float* eps,
                          float* stats, hipStream_t s) {
  const RenderDims d{ll, m, xx, yy, zz, simple};
  float* tab = stats + (size_t)12 * ll * xx * yy;
  hipLaunchKernelGGL(k_render_stats, dim3(blocks_for(3LL * ll * xx * yy)), dim3(256), 0, s, d,
                     layers, stats);
  hipLaunchKernelGGL(k_render_shares, dim3(blocks_for(2LL * ll * zz)), dim3(256), 0, s, d, pos,
                     gs, ge, tab);
  hipLaunchKernelGGL(k_render_combine, dim3(blocks_for(3LL * xx * yy * zz)), dim3(256), 0, s, d,
                     stats, tab, eps);
  return hipGetLastError();
}

hipError_t render_backward(int ll, int m, int xx, int yy, int zz, int simple, const float* layers,
                           const float* pos, const float* gs, const float* ge,
                           const float* stats, const float* g_eps, float* g_layers, float* g_pos,
                           char* ws, hipStream_t s) {
  const RenderDims d{ll, m, xx, yy, zz, simple};
  const float* tab = stats + (size_t)12 * ll * xx * yy;
  float* gst = (float*)ws;
  float* part = (float*)(ws + align256(sizeof(float) * (size_t)12 * ll * xx * yy));
  float* gus = (float*)((char*)part + align256(sizeof(float) * (size_t)2 *
                                               render_chunks_host(xx, yy) * 3 * ll * zz));
  hipLaunchKernelGGL(k_render_bwd_stats, dim3(blocks_for(3LL * xx * yy)), dim3(256), 0, s, d,
                     stats, tab, g_eps, gst);
  hipLaunchKernelGGL(k_render_bwd_layers, dim3(blocks_for((long long)ll * 4 * m * m * xx * yy)),
                     dim3(256), 0, s, d, layers, gst, g_layers);
  if (ll > 1 && g_pos) {
    hipLaunchKernelGGL(k_render_bwd_share, dim3(zz, 3, render_chunks_host(xx, yy)), dim3(256),
                       0, s, d, stats, tab, g_eps, part);
    hipLaunchKernelGGL(k_render_sum_parts, dim3(blocks_for(6LL * ll * zz)), dim3(256), 0, s, d,
                       part, gus);
    hipLaunchKernelGGL(k_render_bwd_pos, dim3(1), dim3(64), 0, s, d, pos, gs, ge, gus, g_pos);
  }
  return hipGetLastError();
}

}  // namespace pjz
