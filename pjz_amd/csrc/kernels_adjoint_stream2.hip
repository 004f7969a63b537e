// Two reversed steps per launch: temporal blocking of depth 2 for the
// adjoint, the counterpart of the two-step forward stepper.  Level 1 reverses
// step n (D^n) at plane x while level 2 reverses step n-1 (D^(n-1)) one plane
// behind at x-1, each with the row/plane dependencies of the one-step
// streaming reverse kernel (kernels_adjoint_stream.hip):
//   A: level-1 lH(x)   needs lK = cb lE at x+1 (prefetch) and y+1 (LDS sK1,
//                      published one plane early from the prefetch)
//   B: level-1 lE(x)   needs level-1 lH at y-1 (LDS sH1) and x-1 (registers)
//      level-2 lH(x-1) needs level-1 lK at x (registers) and y+1 (LDS sK2,
//                      published one plane earlier)
//   C: level-2 lE(x-1) needs level-2 lH at y-1 (LDS sH2) and x-2 (registers)
// Rows y0-2..y0+WY (WY+3 waves): level-1 lH on all, level-1 lE from y0-1,
// level-2 lH on y0-1..y0+WY-1, level-2 lE (stored) on the owned rows; planes
// xa-2, xa-1 are ghost planes.  lE, lH and both psi cotangents ping-pong
// between the two adjoint sets; G is read and written only by the owner:
//   G += lE_in D^n + lE_1 D^(n-1)
// (lE_in the cotangent entering the pair, lE_1 after level 1).
//
// Per cell and pair: read lE, cb, lH, G, D^n, D^(n-1), write lE, lH, G
// (108 B, + psi cotangents on PML cells) against 192 B for two one-step
// launches.
#include <algorithm>
#include <vector>

#include "stream_common.h"

namespace pjz {

struct Adj2Args {
  const float* lEin;
  const float* lHin;
  float* lEout;
  float* lHout;
  const float* luin;  // psi_H cotangent pair (x at +0, y at +P)
  float* luout;
  const float* lpin;  // psi_E cotangent pair
  float* lpout;
  float* G;
  const float* D1;  // D^n (level 1)
  const float* D2;  // D^(n-1) (level 2)
  const float* cb;
  const float* zt;
  const float* mask;
  int X, Y, Z, lo, hi, NP, N, P;
  float dt;
  int pair;  // CPL = 2 with even PML widths and Z: psi entries in aligned pairs
  int LX, nty, ntiles;
  int nsteps;  // steps of the simulation (batched runs)
};

template <int CPL>
struct AdjCoef {
  float ikE[CPL], ikEp[CPL], bE[CPL], cE[CPL], ikH[CPL], ikHm[CPL], bH[CPL], cH[CPL];
  bool pml[CPL];
  int pidx[CPL];
};

// Adjoint of the H half step for one row: lH_out, psi_H cotangent.
//   K: lK at x; Kx: lK y, z at x+1; Ky: lK x, z at y+1 (same row's cells).
template <int CPL>
__device__ inline void adj_h_row(float (&h)[3][CPL], float (&lun)[2][CPL],
                                 const float (&K)[3][CPL], const float (&Kx)[2][CPL],
                                 const float (&Ky)[2][CPL], const float (&lh)[3][CPL],
                                 const float (&lu)[2][CPL], const float (&lp)[2][CPL],
                                 const AdjCoef<CPL>& zc, int zb, int Z, float dt, bool valid) {
  float kxz[CPL], kyz[CPL], vx[CPL], vy[CPL], vxz[CPL], vyz[CPL];
  z_next<CPL>(kxz, K[0], zb, Z);
  z_next<CPL>(kyz, K[1], zb, Z);
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    vx[c] = zc.pml[c] ? zc.cE[c] * (lp[0][c] - K[0][c]) : 0.0f;
    vy[c] = zc.pml[c] ? zc.cE[c] * (lp[1][c] + K[1][c]) : 0.0f;
  }
  z_next<CPL>(vxz, vx, zb, Z);
  z_next<CPL>(vyz, vy, zb, Z);
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    h[0][c] = lh[0][c] - (zc.ikEp[c] * kyz[c] - zc.ikE[c] * K[1][c]) + (Ky[1][c] - K[2][c]) -
              (vyz[c] - vy[c]);
    h[1][c] = lh[1][c] + (zc.ikEp[c] * kxz[c] - zc.ikE[c] * K[0][c]) - (Kx[1][c] - K[2][c]) -
              (vxz[c] - vx[c]);
    h[2][c] = lh[2][c] - (Ky[0][c] - K[0][c]) + (Kx[0][c] - K[1][c]);
    if (!valid) h[0][c] = h[1][c] = h[2][c] = 0.0f;
    lun[0][c] = zc.pml[c] ? zc.bH[c] * lu[0][c] + dt * h[0][c] : 0.0f;
    lun[1][c] = zc.pml[c] ? zc.bH[c] * lu[1][c] - dt * h[1][c] : 0.0f;
  }
}

// Adjoint of the E half step for one row: lE_out, psi_E cotangent.
//   hdn: lH x, z at y-1; hp: lH y, z at x-1; le: lE entering the step.
template <int CPL>
__device__ inline void adj_e_row(float (&len)[3][CPL], float (&lpn)[2][CPL],
                                 const float (&h)[3][CPL], const float (&hdn)[2][CPL],
                                 const float (&hp)[2][CPL], const float (&lun)[2][CPL],
                                 const float (&le)[3][CPL], const float (&cb)[3][CPL],
                                 const float (&sig)[3], const float (&lp)[2][CPL],
                                 const AdjCoef<CPL>& zc, int lane, float dt) {
  float hxz[CPL], hyz[CPL], ux[CPL], uy[CPL], uxz[CPL], uyz[CPL];
  z_prev<CPL>(hxz, h[0], lane);
  z_prev<CPL>(hyz, h[1], lane);
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    ux[c] = zc.pml[c] ? zc.cH[c] * lun[0][c] : 0.0f;
    uy[c] = zc.pml[c] ? zc.cH[c] * lun[1][c] : 0.0f;
  }
  z_prev<CPL>(uxz, ux, lane);
  z_prev<CPL>(uyz, uy, lane);
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    const float gx = dt * ((zc.ikH[c] * h[1][c] - zc.ikHm[c] * hyz[c]) - (h[2][c] - hdn[1][c])) -
                     (uy[c] - uyz[c]);
    const float gy = dt * (-(zc.ikH[c] * h[0][c] - zc.ikHm[c] * hxz[c]) + (h[2][c] - hp[1][c])) -
                     (ux[c] - uxz[c]);
    const float gz = dt * ((h[0][c] - hdn[0][c]) - (h[1][c] - hp[0][c]));
    len[0][c] = (1.0f - sig[0] * cb[0][c]) * le[0][c] + gx;
    len[1][c] = (1.0f - sig[1] * cb[1][c]) * le[1][c] + gy;
    len[2][c] = (1.0f - sig[2] * cb[2][c]) * le[2][c] + gz;
    lpn[0][c] = zc.pml[c] ? zc.bE[c] * (lp[0][c] - cb[0][c] * le[0][c]) : 0.0f;
    lpn[1][c] = zc.pml[c] ? zc.bE[c] * (lp[1][c] + cb[1][c] * le[1][c]) : 0.0f;
  }
}

// Level-1 inputs of one row, fetched one plane ahead.
template <int CPL>
struct Adj2Plane {
  float le[3][CPL];   // lE at x + 1
  float cb[3][CPL];   // cb at x + 1
  float lh[3][CPL];   // lH at x
  float lu[2][CPL];   // psi_H cotangent at x
  float lp[2][CPL];   // psi_E cotangent at x
  float kue[2][CPL];  // lE x, z at (x, y+1) (top row only)
  float kuc[2][CPL];  // cb x, z at (x, y+1): lK = kuc kue, formed where used so
                      // the fetch does not wait for its own loads
  float sig[3];       // absorber at (x, y)
  // Owned rows, with DG: D^n at x (level 1), G and D^(n-1) at x-1 (level
  // 2), fetched with the rest one plane ahead, so the stores of phase C wait
  // for no load.  (Loaded in the plane itself, every store waits for the
  // whole prefetch too: the loads' count is path-dependent, so the wait is
  // vmcnt(0).)  Unused without DG.
  float d1[3][CPL], g[3][CPL], d2[3][CPL];
};

// Carried from plane x-1 to plane x.
template <int CPL>
struct Adj2Carry {
  float le0[3][CPL];  // lE entering the pair at x
  float cb0[3][CPL];  // cb at x
  float e1m[3][CPL];  // level-1 lE at x-1
  float cbm[3][CPL];  // cb at x-1
  float h1m[3][CPL];  // level-1 lH at x-1
  float lu1m[2][CPL]; // level-1 psi_H cotangent at x-1
  float lp1m[2][CPL]; // level-1 psi_E cotangent at x-1
  float h2p[2][CPL];  // level-2 lH y, z at x-2
  float gpm[3][CPL];  // lE_in D^n at x-1 (owned rows)
  float sigm[3];
};

// D^n at x where lvl1, G and D^(n-1) at x-1 where lvl2 (owned rows).
template <int CPL>
__device__ inline void adj2_load_dg(float (&d1)[3][CPL], float (&g)[3][CPL], float (&d2)[3][CPL],
                                    const Adj2Args& a, int x, int yc, int zb, bool lvl1,
                                    bool lvl2) {
  const int N = a.N, Z = a.Z;
  const int o = (x * a.Y + yc) * Z, om = o - a.Y * Z;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    load_col<CPL, true>(d1[k], a.D1 + (k * N + o), zb, Z, lvl1);
    load_col<CPL, true>(g[k], a.G + (k * N + om), zb, Z, lvl2);
    load_col<CPL, true>(d2[k], a.D2 + (k * N + om), zb, Z, lvl2);
  }
}

// DG: also fetch the owned rows' D and G (one plane ahead).  That takes 27
// more registers a lane at two cells a lane; the batch kernel, whose one-cell
// form would then pass 128 registers and lose a quarter of its waves, loads
// them in the plane instead.
template <int CPL, bool DG>
__device__ inline void adj2_fetch(Adj2Plane<CPL>& P, const Adj2Args& a, int x, int yc, bool top,
                                  int zb, bool active, const AdjCoef<CPL>& zc, bool own,
                                  bool row_valid, int xa) {
  const int N = a.N, Y = a.Y, Z = a.Z;
  const int o = (x * Y + yc) * Z;
  const bool nx = active && x + 1 < a.X;
  if (DG)
    adj2_load_dg<CPL>(P.d1, P.g, P.d2, a, x, yc, zb, own && active,
                      own && row_valid && x >= xa && x >= 1);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    load_col<CPL, true>(P.le[k], a.lEin + (k * N + o + Y * Z), zb, Z, nx);
    load_col<CPL, true>(P.cb[k], a.cb + (k * N + o + Y * Z), zb, Z, nx);
    load_col<CPL, true>(P.lh[k], a.lHin + (k * N + o), zb, Z, active);
  }
  if (top) {
    const bool v = active && yc + 1 < Y;
    load_col<CPL, true>(P.kue[0], a.lEin + (o + Z), zb, Z, v);
    load_col<CPL, true>(P.kue[1], a.lEin + (2 * N + o + Z), zb, Z, v);
    load_col<CPL, true>(P.kuc[0], a.cb + (o + Z), zb, Z, v);
    load_col<CPL, true>(P.kuc[1], a.cb + (2 * N + o + Z), zb, Z, v);
  }
  const int pb = (x * Y + yc) * a.NP;
  if (CPL == 2 && a.pair) {
    // A lane's two cells are both PML or both not, with consecutive, 8-byte
    // aligned psi entries: one 64-bit load per component.
    const bool ph = zc.pml[0] && active;
    const int pi = pb + zc.pidx[0];
    float2 u0 = make_float2(0.0f, 0.0f), u1 = u0, p0 = u0, p1 = u0;
    if (ph) {
      u0 = *reinterpret_cast<const float2*>(a.luin + pi);
      u1 = *reinterpret_cast<const float2*>(a.luin + (a.P + pi));
      p0 = *reinterpret_cast<const float2*>(a.lpin + pi);
      p1 = *reinterpret_cast<const float2*>(a.lpin + (a.P + pi));
    }
    P.lu[0][0] = u0.x;
    P.lu[0][CPL - 1] = u0.y;
    P.lu[1][0] = u1.x;
    P.lu[1][CPL - 1] = u1.y;
    P.lp[0][0] = p0.x;
    P.lp[0][CPL - 1] = p0.y;
    P.lp[1][0] = p1.x;
    P.lp[1][CPL - 1] = p1.y;
  } else {
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      const bool ph = zc.pml[c] && active;
      const int pi = pb + zc.pidx[c];
      P.lu[0][c] = ph ? a.luin[pi] : 0.0f;
      P.lu[1][c] = ph ? a.luin[a.P + pi] : 0.0f;
      P.lp[0][c] = ph ? a.lpin[pi] : 0.0f;
      P.lp[1][c] = ph ? a.lpin[a.P + pi] : 0.0f;
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k)
    P.sig[k] = active ? uniform_load(a.mask, (k * a.X + x) * a.Y + yc) : 0.0f;
}

// Row-exchange buffers (cell c of lane l at [c * 64 + l]).
template <int CPL, int WY>
struct Adj2Lds {
  float sK1[2][WY + 3][2][64 * CPL];  // level-1 lK x, z (slot = plane parity)
  float sH1[WY + 3][2][64 * CPL];     // level-1 lH x, z of plane x
  float sK2[2][WY + 3][2][64 * CPL];  // level-1 lE turned lK for level 2, per plane parity
  float sH2[WY + 3][2][64 * CPL];     // level-2 lH x, z of plane x-1
};

template <int CPL, int WY, bool DG>
__device__ __forceinline__ void adj2_plane(const Adj2Args& a, Adj2Lds<CPL, WY>& L,
                                           Adj2Carry<CPL>& s, const Adj2Plane<CPL>& cur,
                                           Adj2Plane<CPL>& nxt, const AdjCoef<CPL>& zc, int x,
                                           int xa, int xb, int w, int y, int yc, int lane,
                                           int zb, bool row_valid, int PAR) {
  constexpr int NW = WY + 3;
  const bool e_rows = w >= 1;
  const bool h2_rows = w >= 1 && w <= WY + 1;
  const bool own = w >= 2 && w <= WY + 1;
  const bool l1 = x < a.X;
  const bool l2 = x >= xa && x >= 1;
  const bool act = row_valid && l1;
  const int N = a.N, Y = a.Y, Z = a.Z;
  const float dt = a.dt;
  // This plane's D^n (level 1 at x) and G, D^(n-1) (level 2 at x-1), owned
  // rows only, used in phases B and C.  Without DG they are requested before
  // the next plane's inputs: waiting for them (in-order completion) then
  // leaves the prefetch in flight.
  float d1[3][CPL], g[3][CPL], d2[3][CPL];
  if (DG) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        d1[k][c] = cur.d1[k][c];
        g[k][c] = cur.g[k][c];
        d2[k][c] = cur.d2[k][c];
      }
  } else {
    adj2_load_dg<CPL>(d1, g, d2, a, x, yc, zb, own && act, own && row_valid && l2);
  }
  if (x + 1 <= xb)
    adj2_fetch<CPL, DG>(nxt, a, x + 1, yc, w == NW - 1, zb, row_valid && x + 1 < a.X, zc, own,
                        row_valid, xa);

  // ---- A: level-1 lH at plane x
  float K1[3][CPL];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int c = 0; c < CPL; ++c) K1[k][c] = s.cb0[k][c] * s.le0[k][c];
  float Kx[2][CPL];
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    Kx[0][c] = cur.cb[1][c] * cur.le[1][c];
    Kx[1][c] = cur.cb[2][c] * cur.le[2][c];
    // lK x, z of plane x+1 for the row below, read next plane.
    L.sK1[PAR ^ 1][w][0][c * 64 + lane] = cur.cb[0][c] * cur.le[0][c];
    L.sK1[PAR ^ 1][w][1][c * 64 + lane] = Kx[1][c];
  }
  float Ky[2][CPL];
  if (w < NW - 1) {
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      Ky[0][c] = L.sK1[PAR][w + 1][0][c * 64 + lane];
      Ky[1][c] = L.sK1[PAR][w + 1][1][c * 64 + lane];
    }
  } else {
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      Ky[0][c] = cur.kuc[0][c] * cur.kue[0][c];
      Ky[1][c] = cur.kuc[1][c] * cur.kue[1][c];
    }
  }
  float h1[3][CPL], lu1[2][CPL];
  adj_h_row<CPL>(h1, lu1, K1, Kx, Ky, cur.lh, cur.lu, cur.lp, zc, zb, Z, dt, act);
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    L.sH1[w][0][c * 64 + lane] = h1[0][c];
    L.sH1[w][1][c * 64 + lane] = h1[2][c];
  }
  lds_barrier();

  // ---- B: level-1 lE at plane x, level-2 lH at plane x-1
  float e1[3][CPL], lp1[2][CPL];
  if (e_rows) {
    float hdn[2][CPL], hp[2][CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      hdn[0][c] = L.sH1[w - 1][0][c * 64 + lane];
      hdn[1][c] = L.sH1[w - 1][1][c * 64 + lane];
      hp[0][c] = s.h1m[1][c];
      hp[1][c] = s.h1m[2][c];
    }
    adj_e_row<CPL>(e1, lp1, h1, hdn, hp, lu1, s.le0, s.cb0, cur.sig, cur.lp, zc, lane, dt);
    if (!act) {
#pragma unroll
      for (int c = 0; c < CPL; ++c) e1[0][c] = e1[1][c] = e1[2][c] = lp1[0][c] = lp1[1][c] = 0.0f;
    }
  } else {
#pragma unroll
    for (int c = 0; c < CPL; ++c) e1[0][c] = e1[1][c] = e1[2][c] = lp1[0][c] = lp1[1][c] = 0.0f;
  }
  // lE_in D^n at x (owned rows), added to G with level 2 next plane.
  float gp[3][CPL];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int c = 0; c < CPL; ++c) gp[k][c] = s.le0[k][c] * d1[k][c];
  float K2x[2][CPL];  // level-2 lK y, z at x (this row)
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    K2x[0][c] = s.cb0[1][c] * e1[1][c];
    K2x[1][c] = s.cb0[2][c] * e1[2][c];
    L.sK2[PAR][w][0][c * 64 + lane] = s.cb0[0][c] * e1[0][c];
    L.sK2[PAR][w][1][c * 64 + lane] = K2x[1][c];
  }
  float h2[3][CPL], lu2[2][CPL];
  if (h2_rows && l2) {
    float K2[3][CPL], K2y[2][CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
#pragma unroll
      for (int k = 0; k < 3; ++k) K2[k][c] = s.cbm[k][c] * s.e1m[k][c];
      K2y[0][c] = L.sK2[PAR ^ 1][w + 1][0][c * 64 + lane];
      K2y[1][c] = L.sK2[PAR ^ 1][w + 1][1][c * 64 + lane];
    }
    adj_h_row<CPL>(h2, lu2, K2, K2x, K2y, s.h1m, s.lu1m, s.lp1m, zc, zb, Z, dt, row_valid);
  } else {
#pragma unroll
    for (int c = 0; c < CPL; ++c) h2[0][c] = h2[1][c] = h2[2][c] = lu2[0][c] = lu2[1][c] = 0.0f;
  }
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    L.sH2[w][0][c * 64 + lane] = h2[0][c];
    L.sH2[w][1][c * 64 + lane] = h2[2][c];
  }
  lds_barrier();

  // ---- C: level-2 lE at plane x-1 (owned rows), stores
  if (own && l2 && row_valid && x - 1 >= xa) {
    float hdn[2][CPL], e2[3][CPL], lp2[2][CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      hdn[0][c] = L.sH2[w - 1][0][c * 64 + lane];
      hdn[1][c] = L.sH2[w - 1][1][c * 64 + lane];
    }
    adj_e_row<CPL>(e2, lp2, h2, hdn, s.h2p, lu2, s.e1m, s.cbm, s.sigm, s.lp1m, zc, lane, dt);
    const int o = ((x - 1) * Y + y) * Z;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float gn[CPL];
#pragma unroll
      for (int c = 0; c < CPL; ++c)
        gn[c] = g[k][c] + s.gpm[k][c] + s.e1m[k][c] * d2[k][c];
      store_col<CPL, true>(a.lEout + (k * N + o), e2[k], zb, Z);
      store_col<CPL, true>(a.lHout + (k * N + o), h2[k], zb, Z);
      store_col<CPL, true>(a.G + (k * N + o), gn, zb, Z);
    }
    const int pb = ((x - 1) * Y + y) * a.NP;
    if (CPL == 2 && a.pair) {
      if (zc.pml[0]) {
        const int pi = pb + zc.pidx[0];
        *reinterpret_cast<float2*>(a.luout + pi) = make_float2(lu2[0][0], lu2[0][CPL - 1]);
        *reinterpret_cast<float2*>(a.luout + a.P + pi) = make_float2(lu2[1][0], lu2[1][CPL - 1]);
        *reinterpret_cast<float2*>(a.lpout + pi) = make_float2(lp2[0][0], lp2[0][CPL - 1]);
        *reinterpret_cast<float2*>(a.lpout + a.P + pi) = make_float2(lp2[1][0], lp2[1][CPL - 1]);
      }
    } else {
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        if (zc.pml[c]) {
          const int pi = pb + zc.pidx[c];
          a.luout[pi] = lu2[0][c];
          a.luout[a.P + pi] = lu2[1][c];
          a.lpout[pi] = lp2[0][c];
          a.lpout[a.P + pi] = lp2[1][c];
        }
      }
    }
  }

  // ---- carry to plane x+1
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    s.h2p[0][c] = h2[1][c];
    s.h2p[1][c] = h2[2][c];
    s.lu1m[0][c] = lu1[0][c];
    s.lu1m[1][c] = lu1[1][c];
    s.lp1m[0][c] = lp1[0][c];
    s.lp1m[1][c] = lp1[1][c];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      s.e1m[k][c] = e1[k][c];
      s.h1m[k][c] = h1[k][c];
      s.cbm[k][c] = s.cb0[k][c];
      s.gpm[k][c] = gp[k][c];
      s.le0[k][c] = cur.le[k][c];
      s.cb0[k][c] = cur.cb[k][c];
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) s.sigm[k] = cur.sig[k];
}

// One tile (WY rows x LX planes) of one pair of reversed steps.
template <int CPL, int WY, bool DG>
__device__ __forceinline__ void adj2_tile(const Adj2Args& a, int tile, Adj2Lds<CPL, WY>& L) {
  constexpr int NW = WY + 3;
  const int Z = a.Z;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int zb = lane * CPL;
  const int ty = tile % a.nty, tx = tile / a.nty;
  const int y = ty * WY - 2 + w;
  const bool row_valid = y >= 0 && y < a.Y;
  const int yc = min(max(y, 0), a.Y - 1);
  const int xa = tx * a.LX;
  const int xb = min(xa + a.LX, a.X);

  AdjCoef<CPL> zc;
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    const int z = zb + c;
    const int zz = min(z, Z - 1);
    zc.ikE[c] = a.zt[IKE * Z + zz];
    zc.ikEp[c] = z + 1 < Z ? a.zt[IKE * Z + z + 1] : 0.0f;
    zc.bE[c] = a.zt[BE * Z + zz];
    zc.cE[c] = a.zt[CE * Z + zz];
    zc.ikH[c] = a.zt[IKH * Z + zz];
    zc.ikHm[c] = z >= 1 && z - 1 < Z ? a.zt[IKH * Z + z - 1] : 0.0f;
    zc.bH[c] = a.zt[BH * Z + zz];
    zc.cH[c] = a.zt[CH * Z + zz];
    zc.pml[c] = z < Z && (z < a.lo || z >= Z - a.hi);
    zc.pidx[c] = z < a.lo ? z : z - (Z - a.hi) + a.lo;
  }

  Adj2Carry<CPL> s;
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
#pragma unroll
    for (int k = 0; k < 3; ++k) s.e1m[k][c] = s.cbm[k][c] = s.h1m[k][c] = s.gpm[k][c] = 0.0f;
    s.lu1m[0][c] = s.lu1m[1][c] = s.lp1m[0][c] = s.lp1m[1][c] = 0.0f;
    s.h2p[0][c] = s.h2p[1][c] = 0.0f;
  }
  s.sigm[0] = s.sigm[1] = s.sigm[2] = 0.0f;
  const int xs = max(xa - 2, 0);
  {
    const int o = (xs * a.Y + yc) * Z;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      load_col<CPL, true>(s.le0[k], a.lEin + (k * a.N + o), zb, Z, row_valid);
      load_col<CPL, true>(s.cb0[k], a.cb + (k * a.N + o), zb, Z, row_valid);
    }
  }
  Adj2Plane<CPL> cur, nxt;
  adj2_fetch<CPL, DG>(cur, a, xs, yc, w == NW - 1, zb, row_valid, zc, w >= 2 && w <= WY + 1,
                      row_valid, xa);
  // lK x, z of the first plane for the row below.
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    L.sK1[0][w][0][c * 64 + lane] = s.cb0[0][c] * s.le0[0][c];
    L.sK1[0][w][1][c * 64 + lane] = s.cb0[2][c] * s.le0[2][c];
  }
  lds_barrier();
  int par = 0;
  for (int x = xs; x <= xb; ++x) {
    adj2_plane<CPL, WY, DG>(a, L, s, cur, nxt, zc, x, xa, xb, w, y, yc, lane, zb, row_valid, par);
    cur = nxt;
    par ^= 1;
  }
}

template <int CPL, int WY>
__global__ __launch_bounds__(64 * (WY + 3)) void k_adj_stream2(Adj2Args a) {
  __shared__ Adj2Lds<CPL, WY> L;
  // Tiles adjacent in y (sharing ghost rows) go to the same XCD.
  int tile = blockIdx.x;
  if ((a.ntiles & 7) == 0) tile = (tile & 7) * (a.ntiles >> 3) + (tile >> 3);
  adj2_tile<CPL, WY, true>(a, tile, L);
}

// Batched pairs of reversed steps: persistent workgroups take (simulation,
// tile) items from a device counter.  args holds two views per simulation
// (which adjoint set is read); a simulation that joined the backward at
// launch lstart[s] reads set (launch - lstart[s]) & 1.  D^step sits at slot
// step - seg_begin of the view's D base, D^(step-1) one slot lower.
template <int CPL, int WY>
__global__ __launch_bounds__(64 * (WY + 3)) void k_adj_stream2_batch(
    const Adj2Args* __restrict__ args, const int* __restrict__ tile_start,
    const int* __restrict__ lstart, int nsims, int total, int* __restrict__ counter, int launch,
    int step, int seg_begin) {
  __shared__ Adj2Lds<CPL, WY> L;
  __shared__ int s_item, s_next;
  if (threadIdx.x == 0) s_next = 0;
  for (;;) {
    if (threadIdx.x == 0) s_item = pop_item(counter, total, s_next);
    __syncthreads();
    const int t = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (t >= total) break;
    int lo = 0, hi = nsims - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tile_start[mid] <= t) lo = mid;
      else hi = mid - 1;
    }
    if (step >= args[2 * lo].nsteps) continue;
    Adj2Args a = args[2 * lo + ((launch - lstart[lo]) & 1)];
    const long long f3 = 3LL * a.N;
    a.D1 += (long long)(step - seg_begin) * f3;
    a.D2 = a.D1 - f3;
    adj2_tile<CPL, WY, false>(a, t - tile_start[lo], L);
  }
}

// Whether the two-step reverse kernel handles this problem.
bool adjoint_stream2_supported(const Geom& g) {
  return (g.Z <= 64) || (g.Z <= 128 && g.Z % 2 == 0);
}

// 4 rows per tile (7 waves): the two levels' carried state needs ~200
// registers a lane, which 2 waves per SIMD allow without spills (8 rows,
// 11 waves, would cap it at 168 and spill: measured slower).
static constexpr int kAdj2WY = 4;

static int device_cus2() {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  return cus;
}

// Workgroups of a kernel that fit on one CU at once (occupancy query): at
// one cell per lane the 7-wave reverse kernel needs ~120 registers, so two
// workgroups share a CU; at two cells per lane (~200) one does.
template <typename K>
static int adj2_resident(K kernel) {
  int b = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, reinterpret_cast<const void*>(kernel),
                                                   64 * (kAdj2WY + 3), 0) != hipSuccess || b < 1)
    b = 1;
  return b;
}

template <int CPL>
static int adj2_resident_single() {
  static const int n = adj2_resident(&k_adj_stream2<CPL, kAdj2WY>);
  return n;
}

template <int CPL>
static int adj2_resident_batch() {
  static const int n = adj2_resident(&k_adj_stream2_batch<CPL, kAdj2WY>);
  return n;
}

// Arguments of a pair of reversed steps reading adjoint set acur; LX = 0
// picks x chunks giving about one tile per resident workgroup slot.
static Adj2Args adj2_args(const Geom& g, const pjz_fdtd_problem& p, const float* mask, char* ws,
                          const Layout& L, const float* D1, const float* D2, int acur, int LX) {
  Adj2Args a;
  const int in = acur ? 1 : 0, out = 1 - in;
  a.lEin = (const float*)(ws + L.lE_[in]);
  a.lHin = (const float*)(ws + L.lH_[in]);
  a.luin = (const float*)(ws + L.lu_[in]);
  a.lpin = (const float*)(ws + L.lpE_[in]);
  a.lEout = (float*)(ws + L.lE_[out]);
  a.lHout = (float*)(ws + L.lH_[out]);
  a.luout = (float*)(ws + L.lu_[out]);
  a.lpout = (float*)(ws + L.lpE_[out]);
  a.G = (float*)(ws + L.accG);
  a.D1 = D1;
  a.D2 = D2;
  a.cb = (const float*)(ws + L.cb);
  a.zt = (const float*)(ws + L.zt);
  a.mask = mask;
  a.X = g.X;
  a.Y = g.Y;
  a.Z = g.Z;
  a.lo = g.lo;
  a.hi = g.hi;
  a.NP = g.NP;
  a.N = (int)g.N;
  a.P = g.X * g.Y * g.NP;
  a.dt = p.dt;
  a.pair = g.Z > 64 && g.Z <= 128 && g.Z % 2 == 0 && g.lo % 2 == 0 && g.hi % 2 == 0;
  a.nty = (g.Y + kAdj2WY - 1) / kAdj2WY;
  if (LX <= 0) {
    const int slots =
        device_cus2() * (g.Z <= 64 ? adj2_resident_single<1>() : adj2_resident_single<2>());
    LX = grid_chunk_x(g.X, a.nty, slots, 2);
  }
  a.LX = LX;
  a.ntiles = a.nty * ((g.X + a.LX - 1) / a.LX);
  a.nsteps = steps_of(p);
  return a;
}

hipError_t launch_adjoint_stream2(const Geom& g, const pjz_fdtd_problem& p, const float* mask,
                                  char* ws, const Layout& L, const float* D1, const float* D2,
                                  int acur, hipStream_t s) {
  if (!adjoint_stream2_supported(g)) return hipErrorInvalidValue;
  Adj2Args a = adj2_args(g, p, mask, ws, L, D1, D2, acur, 0);
  constexpr int WY = kAdj2WY;
  if (g.Z <= 64)
    hipLaunchKernelGGL((k_adj_stream2<1, WY>), dim3(a.ntiles), dim3(64 * (WY + 3)), 0, s, a);
  else
    hipLaunchKernelGGL((k_adj_stream2<2, WY>), dim3(a.ntiles), dim3(64 * (WY + 3)), 0, s, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------- batches
// Meta layout: two argument views per simulation (cells-per-lane class
// order) | tile prefix sums (n_c + 1 per class) | lstart (class order) |
// queue counters (2 x kQueues per launch slot of a segment).
static constexpr int kAdj2BatchLX = 128;

size_t adjoint2_batch_meta_bytes(int n, int S) {
  return align256(sizeof(Adj2Args) * 2 * (size_t)n) + align256(sizeof(int) * (size_t)(n + 2)) +
         align256(sizeof(int) * (size_t)n) + align256(sizeof(int) * 2 * kQueues * (size_t)std::max(S, 1));
}

hipError_t adjoint2_batch_prepare(const BatchItem* items, int n, int S, const int* lstart,
                                  char* meta, Adj2BatchPlan& plan, hipStream_t s) {
  std::vector<int> cls[2];
  for (int i = 0; i < n; ++i) {
    if (!adjoint_stream2_supported(items[i].g)) return hipErrorInvalidValue;
    cls[items[i].g.Z <= 64 ? 0 : 1].push_back(i);
  }
  plan.n = n;
  plan.S = std::max(S, 1);
  plan.meta = meta;
  plan.cus = device_cus2();
  plan.args_bytes = align256(sizeof(Adj2Args) * 2 * (size_t)n);
  plan.ts_bytes = align256(sizeof(int) * (size_t)(n + 2));
  plan.ls_bytes = align256(sizeof(int) * (size_t)n);
  std::vector<char> host(plan.args_bytes + plan.ts_bytes + plan.ls_bytes, 0);
  Adj2Args* hargs = (Adj2Args*)host.data();
  int* hts = (int*)(host.data() + plan.args_bytes);
  int* hls = (int*)(host.data() + plan.args_bytes + plan.ts_bytes);
  int off = 0;
  for (int c = 0; c < 2; ++c) {
    plan.base[c] = off;
    plan.count[c] = (int)cls[c].size();
    std::vector<int> xs, ntys;
    for (int i : cls[c]) {
      xs.push_back(items[i].g.X);
      ntys.push_back((items[i].g.Y + kAdj2WY - 1) / kAdj2WY);
    }
    const int slots =
        plan.cus * (c == 0 ? adj2_resident_batch<1>() : adj2_resident_batch<2>());
    const int cLX = batch_chunk_x(xs.data(), ntys.data(), (int)xs.size(), slots, kAdj2BatchLX);
    int t = 0;
    for (size_t k = 0; k < cls[c].size(); ++k) {
      const int i = cls[c][k];
      const BatchItem& it = items[i];
      const int LX = std::min(cLX, it.g.X);
      for (int v = 0; v < 2; ++v)
        hargs[2 * (off + k) + v] =
            adj2_args(it.g, *it.p, it.mask, it.ws, *it.L, it.capture, it.capture, v, LX);
      hts[off + c + k] = t;
      hls[off + k] = lstart[i];
      t += hargs[2 * (off + k)].ntiles;
    }
    hts[off + c + cls[c].size()] = t;
    plan.total[c] = t;
    off += (int)cls[c].size();
  }
  hipError_t e = hipMemcpyAsync(meta, host.data(), host.size(), hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return e;
  // The host image must outlive the copy.
  return hipStreamSynchronize(s);
}

hipError_t adjoint2_batch_begin_segment(const Adj2BatchPlan& plan, hipStream_t s) {
  int* counters = (int*)(plan.meta + plan.args_bytes + plan.ts_bytes + plan.ls_bytes);
  return launch_zero((float*)counters, 2 * kQueues * (size_t)plan.S, s);
}

hipError_t launch_adjoint2_batch_step(const Adj2BatchPlan& plan, int launch, int step,
                                      int seg_begin, hipStream_t s) {
  const int slot = step - seg_begin;
  if (slot < 1 || slot >= plan.S) return hipErrorInvalidValue;
  const Adj2Args* dargs = (const Adj2Args*)plan.meta;
  const int* dts = (const int*)(plan.meta + plan.args_bytes);
  const int* dls = (const int*)(plan.meta + plan.args_bytes + plan.ts_bytes);
  int* counters = (int*)(plan.meta + plan.args_bytes + plan.ts_bytes + plan.ls_bytes);
  constexpr int WY = kAdj2WY;
  for (int c = 0; c < 2; ++c) {
    const int nc = plan.count[c];
    if (nc == 0) continue;
    const int grid = std::min(
        plan.total[c], plan.cus * (c == 0 ? adj2_resident_batch<1>() : adj2_resident_batch<2>()));
    const Adj2Args* a = dargs + 2 * plan.base[c];
    const int* ts = dts + plan.base[c] + c;
    const int* ls = dls + plan.base[c];
    int* ctr = counters + (2 * slot + c) * kQueues;
    if (c == 0)
      hipLaunchKernelGGL((k_adj_stream2_batch<1, WY>), dim3(grid), dim3(64 * (WY + 3)), 0, s, a,
                         ts, ls, nc, plan.total[c], ctr, launch, step, seg_begin);
    else
      hipLaunchKernelGGL((k_adj_stream2_batch<2, WY>), dim3(grid), dim3(64 * (WY + 3)), 0, s, a,
                         ts, ls, nc, plan.total[c], ctr, launch, step, seg_begin);
  }
  return hipGetLastError();
}

}  // namespace pjz
