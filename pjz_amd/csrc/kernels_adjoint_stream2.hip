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

// Absorber values of row yc at plane x where ok (wave-uniform scalar loads;
// the plane index is clamped so that a load issued ahead of its condition
// stays in the array).
__device__ inline void adj2_sig(float (&sig)[3], const Adj2Args& a, int x, int yc, bool ok) {
  const int xm = min(x, a.X - 1);
#pragma unroll
  for (int k = 0; k < 3; ++k) sig[k] = ok ? uniform_load(a.mask, (k * a.X + xm) * a.Y + yc) : 0.0f;
}

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
      gcfloat* lu = to_global(a.luin);
      gcfloat* lp = to_global(a.lpin);
      const f32x2 a0 = *reinterpret_cast<gcf32x2*>(lu + pi);
      const f32x2 a1 = *reinterpret_cast<gcf32x2*>(lu + (a.P + pi));
      const f32x2 b0 = *reinterpret_cast<gcf32x2*>(lp + pi);
      const f32x2 b1 = *reinterpret_cast<gcf32x2*>(lp + (a.P + pi));
      u0 = make_float2(a0.x, a0.y);
      u1 = make_float2(a1.x, a1.y);
      p0 = make_float2(b0.x, b0.y);
      p1 = make_float2(b1.x, b1.y);
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
      gcfloat* lu = to_global(a.luin);
      gcfloat* lp = to_global(a.lpin);
      P.lu[0][c] = ph ? lu[pi] : 0.0f;
      P.lu[1][c] = ph ? lu[a.P + pi] : 0.0f;
      P.lp[0][c] = ph ? lp[pi] : 0.0f;
      P.lp[1][c] = ph ? lp[a.P + pi] : 0.0f;
    }
  }
  adj2_sig(P.sig, a, x, yc, active);
}

// Row-exchange buffers (cell c of lane l at [c * 64 + l]).
template <int CPL, int WY>
struct Adj2Lds {
  float sK1[2][WY + 3][2][64 * CPL];  // level-1 lK x, z (slot = plane parity)
  float sH1[WY + 3][2][64 * CPL];     // level-1 lH x, z of plane x
  float sK2[2][WY + 3][2][64 * CPL];  // level-1 lE turned lK for level 2, per plane parity
  float sH2[WY + 3][2][64 * CPL];     // level-2 lH x, z of plane x-1
};

template <int CPL, int WY, bool DG, bool RING = false>
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
  // leaves the prefetch in flight.  RING: the caller filled cur with every
  // input of the plane, and nothing is fetched here.
  float d1[3][CPL], g[3][CPL], d2[3][CPL];
  if (DG || RING) {
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
  if (!RING && x + 1 <= xb)
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
    gfloat* lu = to_global(a.luout);
    gfloat* lp = to_global(a.lpout);
    if (CPL == 2 && a.pair) {
      if (zc.pml[0]) {
        const int pi = pb + zc.pidx[0];
        *reinterpret_cast<gf32x2*>(lu + pi) = f32x2{lu2[0][0], lu2[0][CPL - 1]};
        *reinterpret_cast<gf32x2*>(lu + a.P + pi) = f32x2{lu2[1][0], lu2[1][CPL - 1]};
        *reinterpret_cast<gf32x2*>(lp + pi) = f32x2{lp2[0][0], lp2[0][CPL - 1]};
        *reinterpret_cast<gf32x2*>(lp + a.P + pi) = f32x2{lp2[1][0], lp2[1][CPL - 1]};
      }
    } else {
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        if (zc.pml[c]) {
          const int pi = pb + zc.pidx[c];
          lu[pi] = lu2[0][c];
          lu[a.P + pi] = lu2[1][c];
          lp[pi] = lp2[0][c];
          lp[a.P + pi] = lp2[1][c];
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

// Per-z coefficients of a lane's cells zb .. zb + CPL - 1 and their index in
// the psi arrays.
template <int CPL>
__device__ inline AdjCoef<CPL> make_adjcoef(const Adj2Args& a, int zb) {
  const int Z = a.Z;
  gcfloat* zt = to_global(a.zt);
  AdjCoef<CPL> zc;
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    const int z = zb + c;
    const int zz = min(z, Z - 1);
    zc.ikE[c] = zt[IKE * Z + zz];
    zc.ikEp[c] = z + 1 < Z ? zt[IKE * Z + z + 1] : 0.0f;
    zc.bE[c] = zt[BE * Z + zz];
    zc.cE[c] = zt[CE * Z + zz];
    zc.ikH[c] = zt[IKH * Z + zz];
    zc.ikHm[c] = z >= 1 && z - 1 < Z ? zt[IKH * Z + z - 1] : 0.0f;
    zc.bH[c] = zt[BH * Z + zz];
    zc.cH[c] = zt[CH * Z + zz];
    zc.pml[c] = z < Z && (z < a.lo || z >= Z - a.hi);
    zc.pidx[c] = z < a.lo ? z : z - (Z - a.hi) + a.lo;
  }
  return zc;
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

  const AdjCoef<CPL> zc = make_adjcoef<CPL>(a, zb);

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

// ---------------------------------------------------------------- LDS ring
// Two cells per lane with every input of a plane copied straight into LDS
// two planes ahead (global_load_lds), as the forward's k_stream2_ring does:
// lE and cb at x+1 and lH at x for every row, D^n at x and G, D^(n-1) at x-1
// for the owned rows, the psi cotangents, and the top row's lE, cb of the
// row above.  The copies take no registers (no next-plane struct, no D/G
// prefetch), and two planes are in flight per wave instead of one.
//
// LDS: the row exchange (42 KB), the ring 2 x 7 x 9 columns (63 KB), the
// owned rows' D/G ring 2 x 4 x 9 columns (36 KB), psi 7 KB, top row 4 KB:
// 152 KB, one workgroup per CU as before.
//
// The copies are invisible to the compiler, so before reading plane x's
// slot a wave waits until at most K vector memory operations are
// outstanding: the fewest it surely issued after plane x's copies.  Those
// are plane x+1's copies (7 a row; 9 on the top row; 12 on owned rows), plus
// on owned rows 9 stores (lE, lH, G) of each of planes x-2 and x-1 that
// stored (x-1 >= xa + 1; the psi stores are skipped where no lane is PML).
// Nothing else the loop issues is a vector load.
constexpr int kAdjRingPsi = 32;  // psi entries per row the ring holds (lo + hi)

template <int WY>
struct Adj2RingLds {
  static constexpr int NW = WY + 3;
  Adj2Lds<2, WY> xch;          // row exchange, as in k_adj_stream2
  float ring[2][NW][9][128];   // lE x y z and cb x y z at x+1, lH x y z at x
  float dring[2][WY][9][128];  // owned rows: D^n at x, G and D^(n-1) at x-1
  float psr[2][NW][2][64];     // psi_H cotangent x|y, psi_E cotangent x|y
  float topr[2][4][128];       // top row: lE x, z and cb x, z of the row above
};

// Requests plane p's inputs into ring slot `slot` (p <= X): lH, D^n, psi
// and the top row at p, lE and cb at p+1, G and D^(n-1) at p-1, each plane
// clamped into the domain (the reader zeroes what lies outside; at p = X the
// level-2 data of plane X-1 is still needed).  16-byte copies carry two
// columns: lanes 0-31 the even one, lanes 32-63 the odd one (4 cells a lane;
// heights below 128 repeat the last).
template <int WY>
__device__ inline void adj2_ring_copy(Adj2RingLds<WY>& L, const Adj2Args& a, int slot, int w,
                                      int p, int yc, bool top, bool own, int lane) {
  const int N = a.N, Z = a.Z;
  const int pl = min(p, a.X - 1);
  const int oh = (pl * a.Y + yc) * Z;
  const int oe = (min(p + 1, a.X - 1) * a.Y + yc) * Z;
  const int half = lane >> 5;
  const int dz = min((lane & 31) * 4, Z - 4);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int k = 2 * j + half;
    const float* g = k < 3   ? a.lEin + (k * N + oe)
                     : k < 6 ? a.cb + ((k - 3) * N + oe)
                             : a.lHin + ((k - 6) * N + oh);
    // cb x+1 (columns 4 and 5) nontemporal, as D below (read once)
    if (j == 2) dma16nt(g + dz, lds_addr(&L.ring[slot][w][2 * j][0]));
    else dma16(g + dz, lds_addr(&L.ring[slot][w][2 * j][0]));
  }
  if (lane < 32) dma16(a.lHin + (2 * N + oh) + dz, lds_addr(&L.ring[slot][w][8][0]));
  const int e = lane & 31;
  const int pb = (pl * a.Y + yc) * a.NP;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const float* base = q == 0 ? a.luin : a.lpin;
    const float* g = a.NP > 0 ? base + (half * a.P + pb + min(e, a.NP - 1)) : a.lEin;
    dma4(g, lds_addr(&L.psr[slot][w][q][0]));
  }
  if (top) {
    const int ou = (pl * a.Y + min(yc + 1, a.Y - 1)) * Z + (half ? 2 * N : 0);
    dma16(a.lEin + ou + dz, lds_addr(&L.topr[slot][0][0]));
    dma16(a.cb + ou + dz, lds_addr(&L.topr[slot][2][0]));
  }
  if (own) {
    const int om = (min(max(p - 1, 0), a.X - 1) * a.Y + yc) * Z;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = 2 * j + half;
      const float* g = k < 3   ? a.D1 + (k * N + oh)
                       : k < 6 ? a.G + ((k - 3) * N + om)
                               : a.D2 + ((k - 6) * N + om);
      // copies of D^n or D^(n-1) alone nontemporal (grad_256: -0.6%)
      if (j == 0 || j == 3) dma16nt(g + dz, lds_addr(&L.dring[slot][w - 2][2 * j][0]));
      else dma16(g + dz, lds_addr(&L.dring[slot][w - 2][2 * j][0]));
    }
    if (lane < 32) dma16nt(a.D2 + (2 * N + om) + dz, lds_addr(&L.dring[slot][w - 2][8][0]));
  }
}

// Plane x's inputs from ring slot `slot`, zeroed where the register kernel's
// fetch would not load them (absorber values come separately).
template <int WY>
__device__ inline void adj2_ring_read(Adj2Plane<2>& P, const Adj2RingLds<WY>& L,
                                      const Adj2Args& a, int slot, int w, int x, int yc, bool top,
                                      bool own, bool row_valid, int xa, int zb,
                                      const AdjCoef<2>& zc) {
  const bool active = row_valid && x < a.X;
  const bool nx = active && x + 1 < a.X;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const f32x2 te = *reinterpret_cast<const f32x2*>(&L.ring[slot][w][k][zb]);
    const f32x2 tc = *reinterpret_cast<const f32x2*>(&L.ring[slot][w][3 + k][zb]);
    const f32x2 th = *reinterpret_cast<const f32x2*>(&L.ring[slot][w][6 + k][zb]);
    P.le[k][0] = nx ? te.x : 0.0f;
    P.le[k][1] = nx ? te.y : 0.0f;
    P.cb[k][0] = nx ? tc.x : 0.0f;
    P.cb[k][1] = nx ? tc.y : 0.0f;
    P.lh[k][0] = active ? th.x : 0.0f;
    P.lh[k][1] = active ? th.y : 0.0f;
  }
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const bool ok = zc.pml[c] && active;
      const int i = m * 32 + (zc.pml[c] ? zc.pidx[c] : 0);
      P.lu[m][c] = ok ? L.psr[slot][w][0][i] : 0.0f;
      P.lp[m][c] = ok ? L.psr[slot][w][1][i] : 0.0f;
    }
  // Every wave reads (clamped rows) and selects: zeroing fields on one side
  // of a branch lets the compiler merge the two sides' stores into one
  // through a selected address, which puts the struct in scratch memory.
  const bool v = top && active && yc + 1 < a.Y;
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    const f32x2 te = *reinterpret_cast<const f32x2*>(&L.topr[slot][m][zb]);
    const f32x2 tc = *reinterpret_cast<const f32x2*>(&L.topr[slot][2 + m][zb]);
    P.kue[m][0] = v ? te.x : 0.0f;
    P.kue[m][1] = v ? te.y : 0.0f;
    P.kuc[m][0] = v ? tc.x : 0.0f;
    P.kuc[m][1] = v ? tc.y : 0.0f;
  }
  const int wd = own ? w - 2 : 0;
  const bool l1 = own && active;
  const bool l2 = own && row_valid && x >= xa && x >= 1;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const f32x2 t1 = *reinterpret_cast<const f32x2*>(&L.dring[slot][wd][k][zb]);
    const f32x2 tg = *reinterpret_cast<const f32x2*>(&L.dring[slot][wd][3 + k][zb]);
    const f32x2 t2 = *reinterpret_cast<const f32x2*>(&L.dring[slot][wd][6 + k][zb]);
    P.d1[k][0] = l1 ? t1.x : 0.0f;
    P.d1[k][1] = l1 ? t1.y : 0.0f;
    P.g[k][0] = l2 ? tg.x : 0.0f;
    P.g[k][1] = l2 ? tg.y : 0.0f;
    P.d2[k][0] = l2 ? t2.x : 0.0f;
    P.d2[k][1] = l2 ? t2.y : 0.0f;
  }
}

template <int WY>
__device__ __forceinline__ void adj2_ring_tile(const Adj2Args& a, int tile, Adj2RingLds<WY>& L) {
  constexpr int NW = WY + 3;
  const int Z = a.Z;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int zb = lane * 2;
  const int ty = tile % a.nty, tx = tile / a.nty;
  const int y = ty * WY - 2 + w;
  const bool row_valid = y >= 0 && y < a.Y;
  const int yc = min(max(y, 0), a.Y - 1);
  const int xa = tx * a.LX;
  const int xb = min(xa + a.LX, a.X);
  const bool top = w == NW - 1;
  const bool own = w >= 2 && w <= WY + 1;
  const AdjCoef<2> zc = make_adjcoef<2>(a, zb);

  Adj2Carry<2> s;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
#pragma unroll
    for (int k = 0; k < 3; ++k) s.e1m[k][c] = s.cbm[k][c] = s.h1m[k][c] = s.gpm[k][c] = 0.0f;
    s.lu1m[0][c] = s.lu1m[1][c] = s.lp1m[0][c] = s.lp1m[1][c] = 0.0f;
    s.h2p[0][c] = s.h2p[1][c] = 0.0f;
  }
  s.sigm[0] = s.sigm[1] = s.sigm[2] = 0.0f;
  const int xs = max(xa - 2, 0);
  // Planes past xb are never read: their copies repeat plane xb's (L2 hits).
  adj2_ring_copy<WY>(L, a, 0, w, xs, yc, top, own, lane);
  adj2_ring_copy<WY>(L, a, 1, w, min(xs + 1, xb), yc, top, own, lane);
  {
    const int o = (xs * a.Y + yc) * Z;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      load_col<2, true>(s.le0[k], a.lEin + (k * a.N + o), zb, Z, row_valid);
      load_col<2, true>(s.cb0[k], a.cb + (k * a.N + o), zb, Z, row_valid);
    }
  }
  // lK x, z of the first plane for the row below.
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    L.xch.sK1[0][w][0][c * 64 + lane] = s.cb0[0][c] * s.le0[0][c];
    L.xch.sK1[0][w][1][c * 64 + lane] = s.cb0[2][c] * s.le0[2][c];
  }
  lds_barrier();
  Adj2Plane<2> cur, unused;
  float nsig[3];
  adj2_sig(cur.sig, a, xs, yc, row_valid);
  int par = 0;
  for (int x = xs; x <= xb; ++x) {
    const int slot = (x - xs) & 1;
    // Wait for plane x's copies (see above for the counts).
    if (own) {
      const int n = row_valid ? (x - 1 >= xa + 1) + (x - 2 >= xa + 1) : 0;
      if (n == 2) asm volatile("s_waitcnt vmcnt(30)" ::: "memory");
      else if (n == 1) asm volatile("s_waitcnt vmcnt(21)" ::: "memory");
      else asm volatile("s_waitcnt vmcnt(12)" ::: "memory");
    } else if (top) {
      asm volatile("s_waitcnt vmcnt(9)" ::: "memory");
    } else {
      asm volatile("s_waitcnt vmcnt(7)" ::: "memory");
    }
    adj2_ring_read<WY>(cur, L, a, slot, w, x, yc, top, own, row_valid, xa, zb, zc);
    // This wave's reads of the slot are done before plane x+2 overwrites it.
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    adj2_ring_copy<WY>(L, a, slot, w, min(x + 2, xb), yc, top, own, lane);
    adj2_sig(nsig, a, x + 1, yc, x + 1 <= xb && row_valid && x + 1 < a.X);
    adj2_plane<2, WY, false, true>(a, L.xch, s, cur, unused, zc, x, xa, xb, w, y, yc, lane, zb,
                                   row_valid, par);
#pragma unroll
    for (int k = 0; k < 3; ++k) cur.sig[k] = nsig[k];
    par ^= 1;
  }
  // The last planes' copies (repeats, never read) land before the ring is
  // reused or the workgroup ends.
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

template <int WY>
__global__ __launch_bounds__(64 * (WY + 3)) void k_adj_stream2_ring(Adj2Args a) {
  __shared__ Adj2RingLds<WY> L;
  int tile = blockIdx.x;
  if ((a.ntiles & 7) == 0) tile = (tile & 7) * (a.ntiles >> 3) + (tile >> 3);
  adj2_ring_tile<WY>(a, tile, L);
}

// Batched form of k_adj_stream2_ring (the items and views of
// k_adj_stream2_batch).
template <int WY>
__global__ __launch_bounds__(64 * (WY + 3)) void k_adj_stream2_ring_batch(
    const Adj2Args* __restrict__ args, const int* __restrict__ tile_start,
    const int* __restrict__ lstart, int nsims, int total, int* __restrict__ counter, int launch,
    int step, int seg_begin) {
  __shared__ Adj2RingLds<WY> L;
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
    adj2_ring_tile<WY>(a, t - tile_start[lo], L);
  }
}

// Whether the LDS-ring form takes this problem: heights in (64, 128] that
// are multiples of 4, with at most 32 PML cells.
static bool adjoint_stream2_ring_supported(const Geom& g) {
  return g.Z > 64 && g.Z <= 128 && g.Z % 4 == 0 && g.NP <= kAdjRingPsi;
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

static int adj2_resident_ring_batch() {
  static const int n = adj2_resident(&k_adj_stream2_ring_batch<kAdj2WY>);
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
  else if (adjoint_stream2_ring_supported(g))
    hipLaunchKernelGGL((k_adj_stream2_ring<WY>), dim3(a.ntiles), dim3(64 * (WY + 3)), 0, s, a);
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
  plan.ring1 = true;
  for (int i = 0; i < n; ++i) {
    if (!adjoint_stream2_supported(items[i].g)) return hipErrorInvalidValue;
    cls[items[i].g.Z <= 64 ? 0 : 1].push_back(i);
    if (items[i].g.Z > 64) plan.ring1 = plan.ring1 && adjoint_stream2_ring_supported(items[i].g);
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
    const int slots = plan.cus * (c == 0 ? adj2_resident_batch<1>()
                                  : plan.ring1 ? adj2_resident_ring_batch()
                                               : adj2_resident_batch<2>());
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
    const int grid = std::min(plan.total[c], plan.cus * (c == 0        ? adj2_resident_batch<1>()
                                                         : plan.ring1 ? adj2_resident_ring_batch()
                                                                      : adj2_resident_batch<2>()));
    const Adj2Args* a = dargs + 2 * plan.base[c];
    const int* ts = dts + plan.base[c] + c;
    const int* ls = dls + plan.base[c];
    int* ctr = counters + (2 * slot + c) * kQueues;
    if (c == 0)
      hipLaunchKernelGGL((k_adj_stream2_batch<1, WY>), dim3(grid), dim3(64 * (WY + 3)), 0, s, a,
                         ts, ls, nc, plan.total[c], ctr, launch, step, seg_begin);
    else if (plan.ring1)
      hipLaunchKernelGGL((k_adj_stream2_ring_batch<WY>), dim3(grid), dim3(64 * (WY + 3)), 0, s,
                         a, ts, ls, nc, plan.total[c], ctr, launch, step, seg_begin);
    else
      hipLaunchKernelGGL((k_adj_stream2_batch<2, WY>), dim3(grid), dim3(64 * (WY + 3)), 0, s, a,
                         ts, ls, nc, plan.total[c], ctr, launch, step, seg_begin);
  }
  return hipGetLastError();
}

}  // namespace pjz
