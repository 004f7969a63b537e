// Pieces shared by the temporally blocked streaming steppers (two steps per
// launch, kernels_stream2.hip; three or four, kernels_stream_deep.hip):
// launch arguments, per-z CPML coefficients, source injection, the H and E
// row updates of one level, the one-plane-ahead fetch of a row's inputs,
// snapshot / capture stores and the buffer-resource output of the last level.
#pragma once

#include "stream_common.h"

namespace pjz {

// Data touched only by launches that write snapshots or capture E.  It lives
// in device memory and is read only inside the (uniform) branches that store,
// so the stepping loop does not keep it in (scarce) scalar registers.  (A
// non-inlined store routine is worse: taking the launch arguments by
// reference moves them to scratch memory, which made the kernel 3.8x slower.)
struct Stream2Rare {
  float* out;     // snapshots (nn, 3, wex, wey, wez) or null
  float* cap;     // E^m - E^(m+1) of the whole domain at cap + (m - cap_begin) * 3N
  int cap_begin;
  int wx0, wy0, wz0, wex, wey, wez;
  int ostart, ostop, ostep;
};

// Snapshot of level l's step (SNAPl), capture of every level's change.
enum Stream2Flags { SNAP1 = 1, SNAP2 = 2, CAPTURE = 4, SNAP3 = 8, SNAP4 = 16 };

struct Stream2Args {
  const float* Ein;    // old level: E, and H at Ein + hoff
  float* Eout;         // new level: E, and H at Eout + hoff
  const float* psiIn;  // old level psi: HX, HY, EX, EY at multiples of P
  float* psiOut;
  const float* cb;
  const float* zt;
  const float* mask;
  const float* sfield;
  const float* swave;
  const Stream2Rare* rare;
  int X, Y, Z, lo, hi, NP, N, hoff, P;
  int saxis, spos, squad;
  float dt;
  int n, flags;  // first step of the pair; Stream2Flags of this launch
  int pair;      // CPL = 2 with even PML widths: psi entries in aligned pairs
  int LX, nty, ntiles;
  int nsteps;  // steps of the simulation (batched runs)
};

// The block read through the constant address space: wave-uniform scalar
// loads, which a wait for vector memory operations does not include (the
// LDS-ring kernel counts those itself, and a vector load here would make
// every capturing plane wait for all of them).
typedef const __attribute__((address_space(4))) Stream2Rare* rare_cptr;

// Snapshot slot of a step in the output schedule, or -1.
__device__ inline int out_slot(const Stream2Rare* rg, int step) {
  const rare_cptr r = (rare_cptr)rg;
  if (!r->out || step < r->ostart || step >= r->ostop || (step - r->ostart) % r->ostep != 0)
    return -1;
  return (step - r->ostart) / r->ostep;
}

template <int CPL>
struct ZCoef {
  float ikE[CPL], bE[CPL], cE[CPL], ikH[CPL], bH[CPL], cH[CPL];
  bool pml[CPL];
  int pidx[CPL];
};

// Per-z coefficients of a lane's cells zb .. zb + CPL - 1 (clamped reads past
// zz; those cells are never PML) and their index in the psi arrays.
template <int CPL>
__device__ inline ZCoef<CPL> make_zcoef(const Stream2Args& a, int zb) {
  const int Z = a.Z;
  gcfloat* zt = to_global(a.zt);
  ZCoef<CPL> zc;
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    const int z = min(zb + c, Z - 1);
    zc.ikE[c] = zt[IKE * Z + z];
    zc.bE[c] = zt[BE * Z + z];
    zc.cE[c] = zt[CE * Z + z];
    zc.ikH[c] = zt[IKH * Z + z];
    zc.bH[c] = zt[BH * Z + z];
    zc.cH[c] = zt[CH * Z + z];
    zc.pml[c] = (zb + c < Z) && (zb + c < a.lo || zb + c >= Z - a.hi);
    zc.pidx[c] = zb + c < a.lo ? zb + c : zb + c - (Z - a.hi) + a.lo;
  }
  return zc;
}

// Adds J(step) to the curl terms of cells (x, y, zb + c).  As in
// add_source, the two source components land in fixed slots on every branch
// (a lane's z cell is picked by select, not by index) and are added with
// selects on the uniform axis, so the curl terms stay in registers.
template <int CPL>
__device__ inline void add_source2(const Stream2Args& a, int step, int x, int y, int zb,
                                   float (&cx)[CPL], float (&cy)[CPL], float (&cz)[CPL]) {
  gcfloat* sfield = to_global(a.sfield);
  gcfloat* swave = to_global(a.swave);
  float ju[CPL], jv[CPL];  // J of (y, z), (x, z) or (x, y) for axis 0, 1, 2
#pragma unroll
  for (int c = 0; c < CPL; ++c) ju[c] = jv[c] = 0.0f;
  if (a.saxis == 0) {
    if (x == a.spos || x == a.spos - 1) {
      const float wv = swave[2 * step + (x == a.spos ? 0 : 1)];
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        if (zb + c < a.Z) {
          const int so = y * a.Z + zb + c;
          ju[c] = sfield[so] * wv;
          jv[c] = sfield[a.Y * a.Z + so] * wv;
        }
      }
    }
  } else if (a.saxis == 1) {
    if (y == a.spos || y == a.spos - 1) {
      const float wv = swave[2 * step + (y == a.spos ? 0 : 1)];
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        if (zb + c < a.Z) {
          const int so = x * a.Z + zb + c;
          ju[c] = sfield[so] * wv;
          jv[c] = sfield[a.X * a.Z + so] * wv;
        }
      }
    }
  } else {
    const int c0 = a.spos - zb;
    if (c0 >= 0 && c0 < CPL) {
      const int xy = x * a.Y + y, XY = a.X * a.Y;
      float jx = sfield[xy] * swave[2 * step];
      float jy = sfield[XY + xy] * swave[2 * step];
      if (a.squad > 1) {
        jx += sfield[2 * XY + xy] * swave[2 * step + 1];
        jy += sfield[3 * XY + xy] * swave[2 * step + 1];
      }
#pragma unroll
      for (int k = 0; k < CPL; ++k) {
        ju[k] = k == c0 ? jx : 0.0f;
        jv[k] = k == c0 ? jy : 0.0f;
      }
    }
  }
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    cx[c] += a.saxis == 0 ? 0.0f : ju[c];
    cy[c] += a.saxis == 0 ? ju[c] : (a.saxis == 1 ? 0.0f : jv[c]);
    cz[c] += a.saxis == 2 ? 0.0f : jv[c];
  }
}

template <int CPL, int W = 64>
__device__ inline void h_row(float (&hn)[3][CPL], float (&psHn)[2][CPL], const float (&e0)[3][CPL],
                             const float (&e1)[3][CPL], const float (&eup)[2][CPL],
                             const float (&hold)[3][CPL], const float (&psH)[2][CPL],
                             const ZCoef<CPL>& zc, int zb, int Z, float dt, bool valid) {
  float exz[CPL], eyz[CPL];
  z_next<CPL, W>(exz, e0[0], zb, Z);
  z_next<CPL, W>(eyz, e0[1], zb, Z);
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    const float dzEy = eyz[c] - e0[1][c];
    const float dzEx = exz[c] - e0[0][c];
    psHn[0][c] = zc.pml[c] ? zc.bH[c] * psH[0][c] + zc.cH[c] * dzEy : 0.0f;
    psHn[1][c] = zc.pml[c] ? zc.bH[c] * psH[1][c] + zc.cH[c] * dzEx : 0.0f;
    const float tEy = zc.ikH[c] * dzEy + psHn[0][c];
    const float tEx = zc.ikH[c] * dzEx + psHn[1][c];
    hn[0][c] = hold[0][c] - dt * ((eup[1][c] - e0[2][c]) - tEy);
    hn[1][c] = hold[1][c] - dt * (tEx - (e1[2][c] - e0[2][c]));
    hn[2][c] = hold[2][c] - dt * ((e1[1][c] - e0[1][c]) - (eup[0][c] - e0[0][c]));
    if (!valid) hn[0][c] = hn[1][c] = hn[2][c] = psHn[0][c] = psHn[1][c] = 0.0f;
  }
}

template <int CPL, int W = 64>
__device__ inline void e_row(float (&en)[3][CPL], float (&psEn)[2][CPL], const float (&e)[3][CPL],
                             const float (&hn)[3][CPL], const float (&hdn)[2][CPL],
                             const float (&hp)[2][CPL], const float (&cb)[3][CPL],
                             const float (&sig)[3], const float (&psE)[2][CPL],
                             const ZCoef<CPL>& zc, const Stream2Args& a, int step, int x, int y,
                             int zb, int lane, bool valid) {
  float hxz[CPL], hyz[CPL];
  z_prev<CPL, W>(hxz, hn[0], lane);
  z_prev<CPL, W>(hyz, hn[1], lane);
  float cx[CPL], cy[CPL], cz[CPL];
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    const float dzHy = hn[1][c] - hyz[c];
    const float dzHx = hn[0][c] - hxz[c];
    psEn[0][c] = zc.pml[c] ? zc.bE[c] * psE[0][c] + zc.cE[c] * dzHy : 0.0f;
    psEn[1][c] = zc.pml[c] ? zc.bE[c] * psE[1][c] + zc.cE[c] * dzHx : 0.0f;
    const float tHy = zc.ikE[c] * dzHy + psEn[0][c];
    const float tHx = zc.ikE[c] * dzHx + psEn[1][c];
    cx[c] = (hn[2][c] - hdn[1][c]) - tHy;
    cy[c] = tHx - (hn[2][c] - hp[1][c]);
    cz[c] = (hn[1][c] - hp[0][c]) - (hn[0][c] - hdn[0][c]);
  }
  // A batched pair may run one step past the simulation's last (odd step
  // counts); the waveform has no row for it and the result is never read.
  if (valid && step < a.nsteps) add_source2<CPL>(a, step, x, y, zb, cx, cy, cz);
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    en[0][c] = (1.0f - sig[0] * cb[0][c]) * e[0][c] + cb[0][c] * cx[c];
    en[1][c] = (1.0f - sig[1] * cb[1][c]) * e[1][c] + cb[1][c] * cy[c];
    en[2][c] = (1.0f - sig[2] * cb[2][c]) * e[2][c] + cb[2][c] * cz[c];
    if (!valid) en[0][c] = en[1][c] = en[2][c] = psEn[0][c] = psEn[1][c] = 0.0f;
  }
}

// State of the two-step kernels carried from plane x-1 to plane x.
template <int CPL>
struct Carry2 {
  float e0[3][CPL];   // input E at x
  float e1m[3][CPL];  // level-1 E at x-1
  float h1m[3][CPL];  // level-1 H at x-1
  float cbm[3][CPL];  // cb at x-1
  float psm[4][CPL];  // level-1 psi at x-1
  float h2p[2][CPL];  // level-2 (Hy, Hz) at x-2
  float sigm[3];
};

// Level-1 inputs of one row at plane x, fetched one plane ahead.
template <int CPL>
struct Plane2 {
  float e[3][CPL];    // input E at x + 1
  float h[3][CPL];    // input H at x
  float cb[3][CPL];   // cb at x (rows with a level-1 E update)
  float ps[4][CPL];   // input psi (HX, HY, EX, EY) at x
  float up[2][CPL];   // input Ex, Ez at (x, y+1) (top row only)
  float sig[3];
};

template <int CPL, bool vec, int RPW = 1>
__device__ inline void fetch2(Plane2<CPL>& P, const Stream2Args& a, int x, int yc, bool top,
                              int zb, bool active, bool e_rows, const ZCoef<CPL>& zc) {
  const int N = a.N, Z = a.Z;
  const int o = (x * a.Y + yc) * Z;
  const int o1 = o + a.Y * Z;
  const bool a1 = active && x + 1 < a.X;
#pragma unroll
  for (int k = 0; k < 3; ++k) load_col<CPL, vec>(P.e[k], a.Ein + (k * N + o1), zb, Z, a1);
#pragma unroll
  for (int k = 0; k < 3; ++k) load_col<CPL, vec>(P.h[k], a.Ein + (a.hoff + k * N + o), zb, Z, active);
  // cb is read by this row's tile and the ghost rows of its y neighbours,
  // which run at the same time; a nontemporal load measured slightly faster.
#pragma unroll
  for (int k = 0; k < 3; ++k)
    load_col<CPL, vec, true>(P.cb[k], a.cb + (k * N + o), zb, Z, active && e_rows);
  if (top) {
    const bool v = active && yc + 1 < a.Y;
    load_col<CPL, vec>(P.up[0], a.Ein + (o + Z), zb, Z, v);
    load_col<CPL, vec>(P.up[1], a.Ein + (2 * N + o + Z), zb, Z, v);
  }
  const int pb = (x * a.Y + yc) * a.NP;
  if (CPL == 2 && a.pair) {
    // Even PML widths and Z: a lane's two cells are both PML or both not,
    // with consecutive, 8-byte aligned psi entries: one 64-bit load each.
    const bool ph = zc.pml[0] && active;
    const int pi = pb + zc.pidx[0];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      f32x2 t = {0.0f, 0.0f};
      if (ph && (k < 2 || e_rows)) t = *reinterpret_cast<gcf32x2*>(to_global(a.psiIn) + (k * a.P + pi));
      P.ps[k][0] = t.x;
      P.ps[k][CPL - 1] = t.y;
    }
  } else {
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      const bool ph = zc.pml[c] && active;
      const int pi = pb + zc.pidx[c];
      gcfloat* ps = to_global(a.psiIn);
      P.ps[0][c] = ph ? ps[pi] : 0.0f;
      P.ps[1][c] = ph ? ps[a.P + pi] : 0.0f;
      P.ps[2][c] = ph && e_rows ? ps[2 * a.P + pi] : 0.0f;
      P.ps[3][c] = ph && e_rows ? ps[3 * a.P + pi] : 0.0f;
    }
  }
  // One row per wave: the absorber value is wave-uniform (scalar load).
  // Two rows per wave: one value per half.  The plane index is clamped so
  // that a scalar load the compiler issues ahead of its condition stays in
  // the array (the deeper kernels fetch up to two planes past the domain).
  const int xm = min(x, a.X - 1);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int mi = (k * a.X + xm) * a.Y + yc;
    if constexpr (RPW == 1)
      P.sig[k] = active && e_rows ? uniform_load(a.mask, mi) : 0.0f;
    else
      P.sig[k] = active && e_rows ? a.mask[mi] : 0.0f;
  }
}

// Writes the snapshot of E after `step` (en) and/or the capture of its change
// over the step (eo - en) for one row at plane x.
template <int CPL, bool vec>
__device__ __forceinline__ void store_rare(const Stream2Args& a, bool snap, bool cap, int step, int x,
                                        int y, int zb, const float (&en)[3][CPL],
                                        const float (&eo)[3][CPL]) {
  const rare_cptr r = (rare_cptr)a.rare;
  if (snap) {
    const int slot = out_slot(a.rare, step);
    if (slot >= 0 && x >= r->wx0 && x < r->wx0 + r->wex && y >= r->wy0 && y < r->wy0 + r->wey) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        gfloat* op = to_global(r->out) + ((((long long)slot * 3 + k) * r->wex + (x - r->wx0)) *
                                              r->wey + (y - r->wy0)) * (long long)r->wez;
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          const int z = zb + c;
          if (z >= r->wz0 && z < r->wz0 + r->wez) op[z - r->wz0] = en[k][c];
        }
      }
    }
  }
  if (cap && r->cap) {
    gfloat* base = to_global(r->cap) + (long long)(step - r->cap_begin) * 3 * a.N;
    const int o = (x * a.Y + y) * a.Z;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float d[CPL];
#pragma unroll
      for (int c = 0; c < CPL; ++c) d[c] = eo[k][c] - en[k][c];
      store_col_global<CPL, vec>(base + (k * a.N + o), d, zb, a.Z);
    }
  }
}

// Level-2 outputs (E, H at Eout + hoff, new psi) written through buffer
// resources: a row or lane with nothing to store uses an out-of-range offset
// (kOff) and its store is dropped, so every plane issues the same stores and
// the wait for the next plane's prefetched loads does not also wait for them.
struct Out2 {
  rsrc_t e, h, psi;  // E and H of the new level: one resource each
  unsigned N4, P4;
};

__device__ inline Out2 make_out2(const Stream2Args& a) {
  Out2 o;
  o.N4 = 4u * (unsigned)a.N;
  o.P4 = 4u * (unsigned)a.P;
  o.e = make_rsrc(a.Eout, 3 * o.N4);
  o.h = make_rsrc(a.Eout + a.hoff, 3 * o.N4);
  o.psi = make_rsrc(a.psiOut, 4 * o.P4);
  return o;
}

// Host side (kernels_stream2.hip).
int device_cus();
// Store resources of the temporally blocked kernels address at most 2^31 bytes.
bool fits_rsrc(const Geom& g);
Stream2Rare rare_of(const pjz_fdtd_problem& p, const OutWin& win, float* capture, int cap_begin);
__global__ void k_put_rare(Stream2Rare* dst, Stream2Rare v);
Stream2Args make_args2(const Geom& g, const pjz_fdtd_problem& p, const float* mask,
                       const Src& src, const OutWin& win, char* ws, const Layout& L, int WY,
                       int LX, int cur);

// Two-step kernel at two cells per lane with its inputs staged through an
// LDS ring (kernels_stream2_ring.hip): the shapes it takes, and one launch.
bool stream2_ring_supported(const Geom& g);
hipError_t launch_stream2_ring(const Stream2Args& a, hipStream_t s);
// Its batched form (8-row tiles): workgroups resident per CU, and one launch
// of `grid` persistent workgroups over the class's items.
int stream2_ring_batch_resident();
void launch_stream2_ring_batch(const Stream2Args* args, const int* tile_start, int nsims,
                               int total, int* counter, const int* step_base, int step_off,
                               int grid, hipStream_t s);

// Batched three-step launches at one cell per lane (kernels_stream_deep.hip):
// rows per tile, workgroups resident per CU, and launch t's persistent grid
// over args (two views per simulation) from step `step`.
int deep_batch_rows();
int deep_batch_resident();
void launch_deep_batch(const Stream2Args* args, const int* tile_start, int nsims, int total,
                       int* counter, int step, int grid, hipStream_t s);

}  // namespace pjz
