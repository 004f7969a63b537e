// Two-step streaming stepper: one launch advances two time steps (n, n+1)
// while reading and writing each field once, i.e. temporal blocking of depth 2
// along the x sweep.
//
// Per workgroup: rows y0-2 .. y0+WY (WY+3 wavefronts, one y row each, whole z
// column per wave as in the one-step kernel).  While sweeping x, level 1
// (step n) works on plane x and level 2 (step n+1) one plane behind, on x-1:
//   A: level-1 H(x)        needs input E at x, x+1 and row y+1 (LDS sE1)
//   B: level-1 E(x)        needs level-1 H at y-1 (LDS sH1) and x-1 (registers)
//      level-2 H(x-1)      needs level-1 E at x-1, x (registers) and row y+1 of
//                          plane x-1 (LDS sE2, published one plane earlier)
//   C: level-2 E(x-1)      needs level-2 H at y-1 (LDS sH2) and x-2 (registers)
// Level 1 runs on rows y0-2..y0+WY (E from y0-1), level 2 H on y0-1..y0+WY-1
// and level 2 E (the stored result) on the owned rows y0..y0+WY-1; planes
// xa-2 and xa-1 are recomputed ghost planes.  Only level-2 results (and the
// snapshots/captures of both levels on owned cells) are written.  E, H and
// psi ping-pong between buffer sets, since ghost rows/planes read cells other
// workgroups rewrite in the same launch.
#include <algorithm>

#include "stream_common.h"

namespace pjz {

struct Stream2Args {
  const float* Ein;
  const float* Hin;
  float* Eout;
  float* Hout;
  const float* psiIn;  // set of 4 psi components (HX, HY, EX, EY), old level
  float* psiOut;       // same, new level
  const float* cb;
  const float* zt;
  const float* mask;
  Geom g;
  Src src;
  OutWin win;
  float dt;
  int n;                  // first of the two steps
  int out1, out2;         // snapshot slots of steps n and n+1, or -1
  float* cap1;            // capture of E after step n (whole domain), or null
  float* cap2;            // capture of E after step n+1, or null
  int LX, nty, ntiles;
};

// Per-lane z coefficients.
template <int CPL>
struct ZCoef {
  float ikE[CPL], bE[CPL], cE[CPL], ikH[CPL], bH[CPL], cH[CPL];
  bool pml[CPL];
  int pidx[CPL];
};

// H half step of one row at one plane.  e0: E at the plane, e1: E at the
// next plane, eup: (Ex, Ez) of the row above, hold: H before, psH: psi_H
// before; writes H and psi_H after.  Zero for rows outside the domain.
template <int CPL>
__device__ inline void h_row(float (&hn)[3][CPL], float (&psHn)[2][CPL], const float (&e0)[3][CPL],
                             const float (&e1)[3][CPL], const float (&eup)[2][CPL],
                             const float (&hold)[3][CPL], const float (&psH)[2][CPL],
                             const ZCoef<CPL>& zc, int zb, int Z, float dt, bool valid) {
  float exz[CPL], eyz[CPL];
  z_next<CPL>(exz, e0[0], zb, Z);
  z_next<CPL>(eyz, e0[1], zb, Z);
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    const float dzEy = eyz[c] - e0[1][c];
    const float dzEx = exz[c] - e0[0][c];
    float tEy = zc.ikH[c] * dzEy, tEx = zc.ikH[c] * dzEx;
    psHn[0][c] = psHn[1][c] = 0.0f;
    if (zc.pml[c]) {
      psHn[0][c] = zc.bH[c] * psH[0][c] + zc.cH[c] * dzEy;
      psHn[1][c] = zc.bH[c] * psH[1][c] + zc.cH[c] * dzEx;
      tEy += psHn[0][c];
      tEx += psHn[1][c];
    }
    hn[0][c] = hold[0][c] - dt * ((eup[1][c] - e0[2][c]) - tEy);
    hn[1][c] = hold[1][c] - dt * (tEx - (e1[2][c] - e0[2][c]));
    hn[2][c] = hold[2][c] - dt * ((e1[1][c] - e0[1][c]) - (eup[0][c] - e0[0][c]));
    if (!valid) hn[0][c] = hn[1][c] = hn[2][c] = psHn[0][c] = psHn[1][c] = 0.0f;
  }
}

// E half step of one row at one plane.  e: E before, hn: new H at the plane,
// hdn: new (Hx, Hz) of the row below, hp: new (Hy, Hz) at the previous plane,
// psE: psi_E before.  Writes E and psi_E after.  Zero for rows outside.
template <int CPL>
__device__ inline void e_row(float (&en)[3][CPL], float (&psEn)[2][CPL], const float (&e)[3][CPL],
                             const float (&hn)[3][CPL], const float (&hdn)[2][CPL],
                             const float (&hp)[2][CPL], const float (&cb)[3][CPL],
                             const float (&sig)[3], const float (&psE)[2][CPL],
                             const ZCoef<CPL>& zc, const Src& src, const Geom& g, int step, int x,
                             int y, int zb, int lane, bool valid) {
  float hxz[CPL], hyz[CPL];
  z_prev<CPL>(hxz, hn[0], lane);
  z_prev<CPL>(hyz, hn[1], lane);
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    const float dzHy = hn[1][c] - hyz[c];
    const float dzHx = hn[0][c] - hxz[c];
    float tHy = zc.ikE[c] * dzHy, tHx = zc.ikE[c] * dzHx;
    psEn[0][c] = psEn[1][c] = 0.0f;
    if (zc.pml[c]) {
      psEn[0][c] = zc.bE[c] * psE[0][c] + zc.cE[c] * dzHy;
      psEn[1][c] = zc.bE[c] * psE[1][c] + zc.cE[c] * dzHx;
      tHy += psEn[0][c];
      tHx += psEn[1][c];
    }
    float cx = (hn[2][c] - hdn[1][c]) - tHy;
    float cy = tHx - (hn[2][c] - hp[1][c]);
    float cz = (hn[1][c] - hp[0][c]) - (hn[0][c] - hdn[0][c]);
    if (valid) add_source(src, g, step, x, y, zb + c, cx, cy, cz);
    en[0][c] = (1.0f - sig[0] * cb[0][c]) * e[0][c] + cb[0][c] * cx;
    en[1][c] = (1.0f - sig[1] * cb[1][c]) * e[1][c] + cb[1][c] * cy;
    en[2][c] = (1.0f - sig[2] * cb[2][c]) * e[2][c] + cb[2][c] * cz;
    if (!valid) en[0][c] = en[1][c] = en[2][c] = psEn[0][c] = psEn[1][c] = 0.0f;
  }
}

// Level-1 inputs of one row at plane x, fetched one plane ahead.
template <int CPL>
struct Plane2 {
  float e[3][CPL];    // input E at x + 1
  float h[3][CPL];    // input H at x
  float cb[3][CPL];   // cb at x (rows with a level-1 E update)
  float psH[2][CPL];  // input psi_H at x
  float psE[2][CPL];  // input psi_E at x
  float up[2][CPL];   // input Ex, Ez at (x, y+1) (top row only)
  float sig[3];
};

template <int CPL, int NW, bool vec>
__device__ inline void fetch2(Plane2<CPL>& P, const Stream2Args& a, int x, int y, int w, int zb,
                              bool active, bool e_rows, const ZCoef<CPL>& zc, long long pstride) {
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
  for (int k = 0; k < 3; ++k)
    load_col<CPL, vec>(P.cb[k], a.cb + k * N + o, zb, Z, active && e_rows);
  if (w == NW - 1) {
    const bool v = active && y + 1 < Y;
    load_col<CPL, vec>(P.up[0], a.Ein + o + Z, zb, Z, v);
    load_col<CPL, vec>(P.up[1], a.Ein + 2 * N + o + Z, zb, Z, v);
  }
  const long long pb = ((long long)x * Y + y) * g.NP;
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    const bool ph = zc.pml[c] && active;
    const long long pi = pb + zc.pidx[c];
    P.psH[0][c] = ph ? a.psiIn[PSI_HX * pstride + pi] : 0.0f;
    P.psH[1][c] = ph ? a.psiIn[PSI_HY * pstride + pi] : 0.0f;
    P.psE[0][c] = ph && e_rows ? a.psiIn[PSI_EX * pstride + pi] : 0.0f;
    P.psE[1][c] = ph && e_rows ? a.psiIn[PSI_EY * pstride + pi] : 0.0f;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k)
    P.sig[k] = active && e_rows ? a.mask[(k * (long long)X + x) * Y + y] : 0.0f;
}

// Stores the E of one row (whole z column) into a snapshot window slot.
template <int CPL>
__device__ inline void store_snapshot(const OutWin& win, int slot, int x, int y, int zb,
                                      const float (&en)[3][CPL]) {
  if (slot < 0 || x < win.x0 || x >= win.x0 + win.ex || y < win.y0 || y >= win.y0 + win.ey)
    return;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float* op = win.out + ((((long long)slot * 3 + k) * win.ex + (x - win.x0)) * win.ey +
                           (y - win.y0)) * win.ez;
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      const int z = zb + c;
      if (z >= win.z0 && z < win.z0 + win.ez) op[z - win.z0] = en[k][c];
    }
  }
}

template <int CPL, int WY, bool vec>
__global__ __launch_bounds__(64 * (WY + 3)) void k_stream2(Stream2Args a) {
  constexpr int NW = WY + 3;
  const Geom g = a.g;
  const int X = g.X, Y = g.Y, Z = g.Z;
  const long long N = g.N;
  const float dt = a.dt;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int zb = lane * CPL;

  const int b = blockIdx.x;
  int tile = b;
  if ((a.ntiles & 7) == 0) tile = (b & 7) * (a.ntiles >> 3) + (b >> 3);
  const int ty = tile % a.nty, tx = tile / a.nty;
  const int y0 = ty * WY;
  const int y = y0 - 2 + w;
  const bool row_valid = y >= 0 && y < Y;
  const bool e_rows = w >= 1;               // rows with a level-1 E update
  const bool h2_rows = w >= 1 && w <= WY + 1;  // rows with a level-2 H update
  const bool own = w >= 2 && w <= WY + 1;   // rows whose level-2 E is the result
  const int xa = tx * a.LX;
  const int xb = min(xa + a.LX, X);

  __shared__ float sE1[NW][2][64 * CPL];     // input Ex, Ez at plane x
  __shared__ float sH1[NW][2][64 * CPL];     // level-1 Hx, Hz at plane x
  __shared__ float sE2[2][NW][2][64 * CPL];  // level-1 Ex, Ez, by plane parity
  __shared__ float sH2[NW][2][64 * CPL];     // level-2 Hx, Hz at plane x-1

  ZCoef<CPL> zc;
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    const int z = min(zb + c, Z - 1);
    zc.ikE[c] = a.zt[IKE * Z + z];
    zc.bE[c] = a.zt[BE * Z + z];
    zc.cE[c] = a.zt[CE * Z + z];
    zc.ikH[c] = a.zt[IKH * Z + z];
    zc.bH[c] = a.zt[BH * Z + z];
    zc.cH[c] = a.zt[CH * Z + z];
    zc.pml[c] = (zb + c < Z) && (zb + c < g.lo || zb + c >= Z - g.hi);
    zc.pidx[c] = zb + c < g.lo ? zb + c : zb + c - (Z - g.hi) + g.lo;
  }
  const long long pstride = (long long)X * Y * g.NP;

  // Carried state: input E at x; level-1 E, H, cb, sigma, psi at x-1;
  // level-2 (Hy, Hz) at x-2.
  float e0[3][CPL], e1m[3][CPL], h1m[3][CPL], cbm[3][CPL], psm[4][CPL], h2p[2][CPL];
  float sigm[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
#pragma unroll
    for (int k = 0; k < 3; ++k) e1m[k][c] = h1m[k][c] = cbm[k][c] = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) psm[k][c] = 0.0f;
    h2p[0][c] = h2p[1][c] = 0.0f;
  }
  const int xs = max(xa - 2, 0);
  {
    const long long o = ((long long)xs * Y + y) * Z;
#pragma unroll
    for (int k = 0; k < 3; ++k) load_col<CPL, vec>(e0[k], a.Ein + k * N + o, zb, Z, row_valid);
  }
  Plane2<CPL> cur, nxt;
  fetch2<CPL, NW, vec>(cur, a, xs, y, w, zb, row_valid, e_rows, zc, pstride);
  int par = 0;

  for (int x = xs; x <= xb; ++x) {
    const bool l1 = x < X;
    const bool l2 = x >= xa && x >= 1;
    const bool st2 = x - 1 >= xa;
    const bool act = row_valid && l1;
    if (x + 1 <= xb)
      fetch2<CPL, NW, vec>(nxt, a, x + 1, y, w, zb, row_valid && x + 1 < X, e_rows, zc, pstride);

    // ---- A: level-1 H at plane x
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      sE1[w][0][zb + c] = e0[0][c];
      sE1[w][1][zb + c] = e0[2][c];
    }
    lds_barrier();
    float eup[2][CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      eup[0][c] = w < NW - 1 ? sE1[w + 1][0][zb + c] : cur.up[0][c];
      eup[1][c] = w < NW - 1 ? sE1[w + 1][1][zb + c] : cur.up[1][c];
    }
    float hn1[3][CPL], psH1[2][CPL];
    h_row<CPL>(hn1, psH1, e0, cur.e, eup, cur.h, cur.psH, zc, zb, Z, dt, act);
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      sH1[w][0][zb + c] = hn1[0][c];
      sH1[w][1][zb + c] = hn1[2][c];
    }
    lds_barrier();

    // ---- B: level-1 E at plane x, then level-2 H at plane x-1
    float en1[3][CPL], psE1[2][CPL];
    if (e_rows) {
      float hdn[2][CPL], hp[2][CPL];
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        hdn[0][c] = sH1[w - 1][0][zb + c];
        hdn[1][c] = sH1[w - 1][1][zb + c];
        hp[0][c] = h1m[1][c];
        hp[1][c] = h1m[2][c];
      }
      e_row<CPL>(en1, psE1, e0, hn1, hdn, hp, cur.cb, cur.sig, cur.psE, zc, a.src, g, a.n, x, y,
                 zb, lane, act);
      if (own && act && x >= xa && x < xb) {
        store_snapshot<CPL>(a.win, a.out1, x, y, zb, en1);
        if (a.cap1) {
          const long long o = ((long long)x * Y + y) * Z;
#pragma unroll
          for (int k = 0; k < 3; ++k) store_col<CPL, vec>(a.cap1 + k * N + o, en1[k], zb, Z);
        }
      }
    } else {
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        en1[0][c] = en1[1][c] = en1[2][c] = 0.0f;
        psE1[0][c] = psE1[1][c] = 0.0f;
      }
    }
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      sE2[par][w][0][zb + c] = en1[0][c];
      sE2[par][w][1][zb + c] = en1[2][c];
    }
    float hn2[3][CPL], psH2[2][CPL];
    if (h2_rows && l2) {
      float e2up[2][CPL], psHm[2][CPL];
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        e2up[0][c] = sE2[par ^ 1][w + 1][0][zb + c];
        e2up[1][c] = sE2[par ^ 1][w + 1][1][zb + c];
        psHm[0][c] = psm[0][c];
        psHm[1][c] = psm[1][c];
      }
      h_row<CPL>(hn2, psH2, e1m, en1, e2up, h1m, psHm, zc, zb, Z, dt, row_valid);
    } else {
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        hn2[0][c] = hn2[1][c] = hn2[2][c] = 0.0f;
        psH2[0][c] = psH2[1][c] = 0.0f;
      }
    }
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      sH2[w][0][zb + c] = hn2[0][c];
      sH2[w][1][zb + c] = hn2[2][c];
    }
    lds_barrier();

    // ---- C: level-2 E at plane x-1 (the result on owned rows)
    if (own && l2 && row_valid) {
      float hdn[2][CPL], psEm[2][CPL], en2[3][CPL], psE2[2][CPL];
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        hdn[0][c] = sH2[w - 1][0][zb + c];
        hdn[1][c] = sH2[w - 1][1][zb + c];
        psEm[0][c] = psm[2][c];
        psEm[1][c] = psm[3][c];
      }
      e_row<CPL>(en2, psE2, e1m, hn2, hdn, h2p, cbm, sigm, psEm, zc, a.src, g, a.n + 1, x - 1,
                 y, zb, lane, true);
      if (st2) {
        const long long o = ((long long)(x - 1) * Y + y) * Z;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          store_col<CPL, vec>(a.Eout + k * N + o, en2[k], zb, Z);
          store_col<CPL, vec>(a.Hout + k * N + o, hn2[k], zb, Z);
        }
        const long long pb = ((long long)(x - 1) * Y + y) * g.NP;
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          if (zc.pml[c]) {
            const long long pi = pb + zc.pidx[c];
            a.psiOut[PSI_HX * pstride + pi] = psH2[0][c];
            a.psiOut[PSI_HY * pstride + pi] = psH2[1][c];
            a.psiOut[PSI_EX * pstride + pi] = psE2[0][c];
            a.psiOut[PSI_EY * pstride + pi] = psE2[1][c];
          }
        }
        store_snapshot<CPL>(a.win, a.out2, x - 1, y, zb, en2);
        if (a.cap2) {
#pragma unroll
          for (int k = 0; k < 3; ++k) store_col<CPL, vec>(a.cap2 + k * N + o, en2[k], zb, Z);
        }
      }
    }

    // ---- rotate
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      h2p[0][c] = hn2[1][c];
      h2p[1][c] = hn2[2][c];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        e1m[k][c] = en1[k][c];
        h1m[k][c] = hn1[k][c];
        cbm[k][c] = cur.cb[k][c];
        e0[k][c] = cur.e[k][c];
      }
      psm[0][c] = psH1[0][c];
      psm[1][c] = psH1[1][c];
      psm[2][c] = psE1[0][c];
      psm[3][c] = psE1[1][c];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) sigm[k] = cur.sig[k];
    cur = nxt;
    par ^= 1;
  }
}

template <int CPL, int WY, bool vec>
static void launch2(Stream2Args& a, hipStream_t s) {
  hipLaunchKernelGGL((k_stream2<CPL, WY, vec>), dim3(a.ntiles), dim3(64 * (WY + 3)), 0, s, a);
}

// Whether the two-step kernel handles this problem (vector loads needed).
bool stream2_supported(const Geom& g) {
  return g.Z <= 256 && ((g.Z <= 64) || (g.Z <= 128 && g.Z % 2 == 0) || (g.Z % 4 == 0));
}

// Runs steps [r.begin, r.end) two at a time; an odd last step uses the
// one-step kernel (same buffer sets).
hipError_t run_stream2(const Geom& g, const pjz_fdtd_problem& p, const float* mask,
                       const Src& src, const OutWin& win, char* ws, const Layout& L,
                       const StepRange& r, hipStream_t s) {
  if (!stream2_supported(g)) return hipErrorInvalidValue;
  const int cpl = g.Z <= 64 ? 1 : (g.Z <= 128 ? 2 : 4);
  const int WY = cpl == 4 ? 4 : (p.tile_y > 0 ? p.tile_y : 8);
  Stream2Args a;
  float* psi = (float*)(ws + L.psi);
  const long long pstride = (long long)g.X * g.Y * g.NP;
  a.cb = (const float*)(ws + L.cb);
  a.zt = (const float*)(ws + L.zt);
  a.mask = mask;
  a.g = g;
  a.src = src;
  a.win = win;
  a.dt = p.dt;
  if (p.tile_x > 0) {
    a.LX = p.tile_x;
  } else {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
      (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int nty = (g.Y + WY - 1) / WY;
    const int ntx = std::max(1, (cus + nty - 1) / nty);
    a.LX = std::max(8, (g.X + ntx - 1) / ntx);
  }
  a.nty = (g.Y + WY - 1) / WY;
  a.ntiles = a.nty * ((g.X + a.LX - 1) / a.LX);
  float* E[2] = {(float*)(ws + L.E0), (float*)(ws + L.E1)};
  float* H[2] = {(float*)(ws + L.H0), (float*)(ws + L.H1)};
  const long long f3 = 3 * g.N;
  auto slot = [&](int step) {
    if (!win.out || step < p.out_start || step >= p.out_stop ||
        (step - p.out_start) % p.out_step != 0)
      return -1;
    return (step - p.out_start) / p.out_step;
  };
  int n = r.begin;
  for (; n + 2 <= r.end; n += 2) {
    const int c = *r.cur;
    a.Ein = E[c];
    a.Hin = H[c];
    a.Eout = E[1 - c];
    a.Hout = H[1 - c];
    a.psiIn = psi + (c ? PSI_SET : 0) * pstride;
    a.psiOut = psi + (c ? 0 : PSI_SET) * pstride;
    a.n = n;
    a.out1 = slot(n);
    a.out2 = slot(n + 1);
    a.cap1 = r.capture ? r.capture + (long long)(n - r.begin + 1) * f3 : nullptr;
    a.cap2 = r.capture ? r.capture + (long long)(n - r.begin + 2) * f3 : nullptr;
    if (cpl == 1) launch2<1, 8, true>(a, s);
    else if (cpl == 4) launch2<4, 4, true>(a, s);
    else if (WY == 4) launch2<2, 4, true>(a, s);
    else if (WY == 12) launch2<2, 12, true>(a, s);
    else if (WY == 13) launch2<2, 13, true>(a, s);
    else launch2<2, 8, true>(a, s);
    *r.cur = 1 - c;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || n >= r.end) return e;
  StepRange tail{n, r.end, r.cur, r.capture ? r.capture + (long long)(n - r.begin) * f3 : nullptr};
  return run_stream(g, p, mask, src, win, ws, L, tail, s);
}

}  // namespace pjz
