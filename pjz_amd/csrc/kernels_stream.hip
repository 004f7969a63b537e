// Streaming stepper: one fused launch per time step (H half step + E half
// step), reading each field once and writing it once.
//
// Mapping (gfx950, wave64):
//   - one wavefront owns one y row of one x plane: the whole z column, CPL
//     consecutive z cells per lane (CPL = 1, 2, 4 for zz <= 64, 128, 256), so
//     z neighbours are one lane away (cross-lane shuffles);
//   - a workgroup owns WY rows [y0, y0+WY) plus one ghost row y0-1 (wave 0)
//     whose H update feeds the E update of row y0; rows exchange E (for H) and
//     the new H (for E) through LDS;
//   - the workgroup sweeps x over a chunk [xa, xb), carrying the new Hy, Hz of
//     plane x-1 in registers; the plane xa-1 is recomputed (ghost plane).
// E, H and psi ping-pong between two buffer sets: neighbouring workgroups
// read the old level on their ghost row/plane while this one writes the new.
#include <algorithm>

#include "stream_common.h"

namespace pjz {

struct StreamArgs {
  const float* Ein;
  const float* Hin;
  float* Eout;
  float* Hout;
  const float* psiHin;  // PSI_HX, PSI_HY of the old level (HY at +pstride)
  float* psiHout;       // same pair, new level
  const float* psiEin;  // PSI_EX, PSI_EY of the old level (EY at +pstride)
  float* psiEout;       // same pair, new level
  const float* cb;
  const float* zt;
  const float* mask;
  Geom g;
  Src src;
  OutWin win;
  float dt;
  int n;        // time step
  int out_idx;  // snapshot slot or -1
  float* Ecap;  // if set, E^n - E^(n+1) of the whole domain is also written here
  int LX;       // x planes per workgroup
  int nty;      // number of y tiles
  int ntiles;   // total tiles
};


// Everything one wave reads for plane x of its row, fetched one plane ahead.
template <int CPL>
struct Plane {
  float e[3][CPL];    // E at x + 1 (E at x is carried from the previous plane)
  float h[3][CPL];    // H at x
  float cb[3][CPL];   // cb at x (owner rows only)
  float psH[2][CPL];  // psi_H pair at x (PML cells)
  float psE[2][CPL];  // psi_E pair at x (PML cells, owner rows only)
  float up[2][CPL];   // Ex, Ez at (x, y+1) (top row of the workgroup only)
  float sig[3];       // absorber at (x, y)
};

template <int CPL, int WY, bool vec>
__device__ inline void fetch_plane(Plane<CPL>& P, const StreamArgs& a, int x, int y, int w,
                                   int zb, bool active, bool owner_cells,
                                   const bool (&pml)[CPL], const int (&pidx)[CPL],
                                   long long pstride) {
  const Geom& g = a.g;
  const long long N = g.N;
  const int X = g.X, Y = g.Y, Z = g.Z;
  const long long o = ((long long)x * Y + y) * Z;
#pragma unroll
  for (int k = 0; k < 3; ++k)
    load_col<CPL, vec>(P.e[k], a.Ein + k * N + o + (long long)Y * Z, zb, Z, active && x + 1 < X);
#pragma unroll
  for (int k = 0; k < 3; ++k) load_col<CPL, vec>(P.h[k], a.Hin + k * N + o, zb, Z, active);
#pragma unroll
  for (int k = 0; k < 3; ++k) load_col<CPL, vec>(P.cb[k], a.cb + k * N + o, zb, Z, owner_cells);
  if (w == WY) {
    const bool v = active && y + 1 < Y;
    load_col<CPL, vec>(P.up[0], a.Ein + o + Z, zb, Z, v);
    load_col<CPL, vec>(P.up[1], a.Ein + 2 * N + o + Z, zb, Z, v);
  }
  const long long pb = ((long long)x * Y + y) * g.NP;
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    const bool ph = pml[c] && active;
    const bool pe = pml[c] && owner_cells;
    P.psH[0][c] = ph ? a.psiHin[pb + pidx[c]] : 0.0f;
    P.psH[1][c] = ph ? a.psiHin[pstride + pb + pidx[c]] : 0.0f;
    P.psE[0][c] = pe ? a.psiEin[pb + pidx[c]] : 0.0f;
    P.psE[1][c] = pe ? a.psiEin[pstride + pb + pidx[c]] : 0.0f;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k)
    P.sig[k] = owner_cells ? a.mask[(k * (long long)X + x) * Y + y] : 0.0f;
}

template <int CPL, int WY, bool vec, int OCC>
__global__ __launch_bounds__(64 * (WY + 1)) __attribute__((amdgpu_waves_per_eu(OCC))) void
k_stream(StreamArgs a) {
  const Geom g = a.g;
  const int X = g.X, Y = g.Y, Z = g.Z;
  const long long N = g.N;
  const float dt = a.dt;
  const int lane = threadIdx.x & 63;
  // Wave index (0 = ghost row).  readfirstlane tells the compiler it is
  // wave-uniform, so row offsets and predicates live in SGPRs and loads use
  // the scalar-base + lane-offset addressing form.
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int zb = lane * CPL;

  // Consecutive workgroup ids go round-robin over the 8 XCDs; remap so that
  // tiles adjacent in y (which share halo rows) run on the same XCD.
  const int b = blockIdx.x;
  int tile = b;
  if ((a.ntiles & 7) == 0) tile = (b & 7) * (a.ntiles >> 3) + (b >> 3);
  const int ty = tile % a.nty, tx = tile / a.nty;
  const int y0 = ty * WY;
  const int y = y0 - 1 + w;
  const bool row_valid = y >= 0 && y < Y;
  const bool owner = w >= 1 && row_valid;
  const int xa = tx * a.LX;
  const int xb = min(xa + a.LX, X);

  // Row exchange (cell c of lane l at [c * 64 + l]: conflict-free banks),
  // double-buffered by plane parity: the input Ex, Ez of plane x+1 are
  // published from the prefetch while plane x is computed, so one barrier
  // per plane (between the H and the E half steps) orders everything.
  __shared__ float sE[2][WY + 1][2][64 * CPL];  // Ex, Ez of a plane, per row
  __shared__ float sH[2][WY + 1][2][64 * CPL];  // new Hx, Hz of a plane, per row

  float ikE[CPL], bE[CPL], cE[CPL], ikH[CPL], bH[CPL], cH[CPL];
  bool pml[CPL];
  int pidx[CPL];
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    int z = min(zb + c, Z - 1);
    ikE[c] = a.zt[IKE * Z + z];
    bE[c] = a.zt[BE * Z + z];
    cE[c] = a.zt[CE * Z + z];
    ikH[c] = a.zt[IKH * Z + z];
    bH[c] = a.zt[BH * Z + z];
    cH[c] = a.zt[CH * Z + z];
    pml[c] = (zb + c < Z) && (zb + c < g.lo || zb + c >= Z - g.hi);
    pidx[c] = zb + c < g.lo ? zb + c : zb + c - (Z - g.hi) + g.lo;
  }
  const long long pstride = (long long)X * Y * g.NP;
  const long long XY = (long long)X * Y;

  float e0[3][CPL], hyp[CPL], hzp[CPL];
  const int xs = max(xa - 1, 0);
  {
    const long long o = ((long long)xs * Y + y) * Z;
#pragma unroll
    for (int k = 0; k < 3; ++k) load_col<CPL, vec>(e0[k], a.Ein + k * N + o, zb, Z, row_valid);
#pragma unroll
    for (int c = 0; c < CPL; ++c) hyp[c] = hzp[c] = 0.0f;
  }
  Plane<CPL> cur, nxt;
  fetch_plane<CPL, WY, vec>(cur, a, xs, y, w, zb, row_valid, owner && xs >= xa, pml, pidx,
                            pstride);
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    sE[0][w][0][c * 64 + lane] = e0[0][c];
    sE[0][w][1][c * 64 + lane] = e0[2][c];
  }
  lds_barrier();

  for (int x = xs; x < xb; ++x) {
    const bool ghost_plane = x < xa;
    const long long o = ((long long)x * Y + y) * Z;
    // Prefetch plane x+1 (zeros past the chunk).
    fetch_plane<CPL, WY, vec>(nxt, a, x + 1, y, w, zb, row_valid && x + 1 < xb,
                              owner && x + 1 < xb, pml, pidx, pstride);
    const int par = (x - xs) & 1;
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      sE[par ^ 1][w][0][c * 64 + lane] = cur.e[0][c];
      sE[par ^ 1][w][1][c * 64 + lane] = cur.e[2][c];
    }
    float eup[2][CPL];
    if (w < WY) {
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        eup[0][c] = sE[par][w + 1][0][c * 64 + lane];
        eup[1][c] = sE[par][w + 1][1][c * 64 + lane];
      }
    } else {
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        eup[0][c] = cur.up[0][c];
        eup[1][c] = cur.up[1][c];
      }
    }

    // ---- H half step at plane x, row y
    float exz[CPL], eyz[CPL], hn[3][CPL];
    z_next<CPL>(exz, e0[0], zb, Z);
    z_next<CPL>(eyz, e0[1], zb, Z);
    const long long pb = ((long long)x * Y + y) * g.NP;
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      float dzEy = eyz[c] - e0[1][c];
      float dzEx = exz[c] - e0[0][c];
      float tEy = ikH[c] * dzEy, tEx = ikH[c] * dzEx;
      if (pml[c]) {
        float p0 = bH[c] * cur.psH[0][c] + cH[c] * dzEy;
        float p1 = bH[c] * cur.psH[1][c] + cH[c] * dzEx;
        if (owner && !ghost_plane) {
          a.psiHout[pb + pidx[c]] = p0;
          a.psiHout[pstride + pb + pidx[c]] = p1;
        }
        tEy += p0;
        tEx += p1;
      }
      hn[0][c] = cur.h[0][c] - dt * ((eup[1][c] - e0[2][c]) - tEy);
      hn[1][c] = cur.h[1][c] - dt * (tEx - (cur.e[2][c] - e0[2][c]));
      hn[2][c] = cur.h[2][c] - dt * ((cur.e[1][c] - e0[1][c]) - (eup[0][c] - e0[0][c]));
      // A row outside the domain (the ghost row y = -1 of the first tile, or
      // rows past Y in the last one) holds no cells: its H stays zero even
      // though the row above it is not.
      if (!row_valid) hn[0][c] = hn[1][c] = hn[2][c] = 0.0f;
    }
    if (owner && !ghost_plane) {
#pragma unroll
      for (int k = 0; k < 3; ++k) store_col<CPL, vec>(a.Hout + k * N + o, hn[k], zb, Z);
    }
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      sH[par][w][0][c * 64 + lane] = hn[0][c];
      sH[par][w][1][c * 64 + lane] = hn[2][c];
    }
    lds_barrier();

    // ---- E half step at plane x, row y
    float hxz[CPL], hyz[CPL];
    z_prev<CPL>(hxz, hn[0], lane);
    z_prev<CPL>(hyz, hn[1], lane);
    if (owner && !ghost_plane) {
      float hdn[2][CPL];
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        hdn[0][c] = sH[par][w - 1][0][c * 64 + lane];
        hdn[1][c] = sH[par][w - 1][1][c * 64 + lane];
      }
      float en[3][CPL];
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        float dzHy = hn[1][c] - hyz[c];
        float dzHx = hn[0][c] - hxz[c];
        float tHy = ikE[c] * dzHy, tHx = ikE[c] * dzHx;
        if (pml[c]) {
          float p0 = bE[c] * cur.psE[0][c] + cE[c] * dzHy;
          float p1 = bE[c] * cur.psE[1][c] + cE[c] * dzHx;
          a.psiEout[pb + pidx[c]] = p0;
          a.psiEout[pstride + pb + pidx[c]] = p1;
          tHy += p0;
          tHx += p1;
        }
        float cx = (hn[2][c] - hdn[1][c]) - tHy;
        float cy = tHx - (hn[2][c] - hzp[c]);
        float cz = (hn[1][c] - hyp[c]) - (hn[0][c] - hdn[0][c]);
        add_source(a.src, g, a.n, x, y, zb + c, cx, cy, cz);
        en[0][c] = (1.0f - cur.sig[0] * cur.cb[0][c]) * e0[0][c] + cur.cb[0][c] * cx;
        en[1][c] = (1.0f - cur.sig[1] * cur.cb[1][c]) * e0[1][c] + cur.cb[1][c] * cy;
        en[2][c] = (1.0f - cur.sig[2] * cur.cb[2][c]) * e0[2][c] + cur.cb[2][c] * cz;
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) store_col<CPL, vec>(a.Eout + k * N + o, en[k], zb, Z);
      if (a.Ecap) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          float d[CPL];
#pragma unroll
          for (int c = 0; c < CPL; ++c) d[c] = e0[k][c] - en[k][c];
          store_col<CPL, vec>(a.Ecap + k * N + o, d, zb, Z);
        }
      }
      if (a.out_idx >= 0 && x >= a.win.x0 && x < a.win.x0 + a.win.ex && y >= a.win.y0 &&
          y < a.win.y0 + a.win.ey) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          float* op = a.win.out +
                      ((((long long)a.out_idx * 3 + k) * a.win.ex + (x - a.win.x0)) * a.win.ey +
                       (y - a.win.y0)) * a.win.ez;
#pragma unroll
          for (int c = 0; c < CPL; ++c) {
            const int z = zb + c;
            if (z >= a.win.z0 && z < a.win.z0 + a.win.ez) op[z - a.win.z0] = en[k][c];
          }
        }
      }
    }
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      hyp[c] = hn[1][c];
      hzp[c] = hn[2][c];
#pragma unroll
      for (int k = 0; k < 3; ++k) e0[k][c] = cur.e[k][c];
    }
    cur = nxt;
  }
}

template <int CPL, int WY, bool vec, int OCC>
static hipError_t launch_stream(StreamArgs& a, const StepRange& r, const pjz_fdtd_problem& p,
                                float* E0, float* H0, float* E1, float* H1, float* psi,
                                long long pstride, hipStream_t s) {
  a.nty = (a.g.Y + WY - 1) / WY;
  a.ntiles = a.nty * ((a.g.X + a.LX - 1) / a.LX);
  const long long f3 = 3 * a.g.N;
  for (int n = r.begin; n < r.end; ++n) {
    a.n = n;
    a.out_idx = -1;
    if (a.win.out && n >= p.out_start && n < p.out_stop && (n - p.out_start) % p.out_step == 0)
      a.out_idx = (n - p.out_start) / p.out_step;
    a.Ecap = r.capture ? r.capture + (long long)(n - r.begin) * f3 : nullptr;
    // The state is read from buffer set *cur and written to the other one.
    const bool even = *r.cur == 0;
    *r.cur = even ? 1 : 0;
    a.Ein = even ? E0 : E1;
    a.Hin = even ? H0 : H1;
    a.Eout = even ? E1 : E0;
    a.Hout = even ? H1 : H0;
    a.psiHin = psi + (even ? 0 : PSI_SET) * pstride;
    a.psiHout = psi + (even ? PSI_SET : 0) * pstride;
    a.psiEin = a.psiHin + PSI_EX * pstride;
    a.psiEout = a.psiHout + PSI_EX * pstride;
    hipLaunchKernelGGL((k_stream<CPL, WY, vec, OCC>), dim3(a.ntiles), dim3(64 * (WY + 1)), 0, s,
                       a);
  }
  return hipGetLastError();
}

hipError_t run_stream(const Geom& g, const pjz_fdtd_problem& p, const float* mask,
                      const Src& src, const OutWin& win, char* ws, const Layout& L,
                      const StepRange& r, hipStream_t s) {
  if (g.Z > 256) return hipErrorInvalidValue;
  StreamArgs a;
  float* psi = (float*)(ws + L.psi);
  const long long pstride = (long long)g.X * g.Y * g.NP;
  a.cb = (const float*)(ws + L.cb);
  a.zt = (const float*)(ws + L.zt);
  a.mask = mask;
  a.g = g;
  a.src = src;
  a.win = win;
  a.dt = p.dt;
  a.Ecap = nullptr;
  // Rows per workgroup: 8 at CPL = 2 (measured best at 512x512x128), else 4.
  const bool cpl2 = g.Z > 64 && g.Z <= 128 && g.Z % 2 == 0;
  int wy = cpl2 ? 8 : 4;
  if (cpl2 && p.tile_y > 0) wy = p.tile_y;
  // x chunks sized so the grid has about one workgroup per CU: the 9-wave
  // workgroups then run in a single round, with 1/LX ghost planes.
  if (p.tile_x > 0) {
    a.LX = p.tile_x;
  } else {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
      (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    a.LX = grid_chunk_x(g.X, (g.Y + wy - 1) / wy, cus, 1);
  }
  float* E0 = (float*)(ws + L.E0);
  float* H0 = (float*)(ws + L.H0);
  float* E1 = (float*)(ws + L.E1);
  float* H1 = (float*)(ws + L.H1);
#define PJZ_LAUNCH(CPL, WY, VEC, OCC) \
  return launch_stream<CPL, WY, VEC, OCC>(a, r, p, E0, H0, E1, H1, psi, pstride, s)
  if (g.Z <= 64) PJZ_LAUNCH(1, 4, true, 1);
  if (g.Z <= 128) {
    if (!cpl2) PJZ_LAUNCH(2, 4, false, 1);
    if (wy == 2) PJZ_LAUNCH(2, 2, true, 1);
    if (wy == 4) PJZ_LAUNCH(2, 4, true, 1);
    if (wy == 12) PJZ_LAUNCH(2, 12, true, 1);
    if (wy == 15) PJZ_LAUNCH(2, 15, true, 1);
    PJZ_LAUNCH(2, 8, true, 1);
  }
  if (g.Z % 4 == 0) PJZ_LAUNCH(4, 4, true, 1);
  PJZ_LAUNCH(4, 4, false, 1);
#undef PJZ_LAUNCH
}

}  // namespace pjz
