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
//
// Instruction economy (the kernel is issue-bound once two steps share a
// pass): offsets are 32-bit (the C ABI bounds 3*cells below 2^31) and H sits
// at a fixed offset from E, so few pointers and no 64-bit index products stay
// live in SGPRs.
#include <algorithm>
#include <functional>
#include <queue>
#include <utility>
#include <vector>

#include "stream2_parts.h"

namespace pjz {

// Row exchange buffers; a row's cell c of lane l (of the LPR = 64 / RPW
// lanes that hold the row) sits at [c * LPR + l], so the lanes of an access
// hit consecutive banks.
template <int CPL, int WY, int RPW = 1>
struct Lds2 {
  static constexpr int LPR = 64 / RPW;
  float sE1[2][WY + 3][2][LPR * CPL];
  float sH1[WY + 3][2][LPR * CPL];
  float sE2[2][WY + 3][2][LPR * CPL];
  float sH2[WY + 3][2][LPR * CPL];
};

// One plane of the sweep: level 1 at x, level 2 at x-1.  PAR selects the
// sE2 slot written this plane (the other slot holds plane x-1) and the sE1
// slot read this plane (input Ex, Ez of plane x, published one plane early
// from the prefetched registers, so level-1 H needs no barrier of its own).
// `w` is the row within the tile (rows y0-2 .. y0+WY) and `lane` the lane
// within the row's LPR lanes; with two rows per wave (RPW = 2) both vary
// within a wave.
template <int CPL, int WY, bool vec, int RPW = 1>
__device__ __forceinline__ void plane2(const Stream2Args& a, const Out2& O,
                                       Lds2<CPL, WY, RPW>& L, Carry2<CPL>& s,
                                       const Plane2<CPL>& cur, Plane2<CPL>& nxt,
                                       const ZCoef<CPL>& zc, int x, int xa, int xb, int w,
                                       int y, int yc, int lane, int zb, bool row_valid,
                                       bool lane_ok, int PAR) {
  constexpr int NW = WY + 3;
  constexpr int LPR = 64 / RPW;
  const bool e_rows = w >= 1;
  const bool h2_rows = w >= 1 && w <= WY + 1;
  const bool own = w >= 2 && w <= WY + 1;
  const bool l1 = x < a.X;
  const bool l2 = x >= xa && x >= 1;
  const bool act = row_valid && l1;
  const int Z = a.Z, N = a.N;
  if (x + 1 <= xb)
    fetch2<CPL, vec, RPW>(nxt, a, x + 1, yc, w == NW - 1, zb, row_valid && x + 1 < a.X, e_rows,
                          zc);

  // ---- A: level-1 H at plane x (input E of the row above from sE1[PAR];
  // this row's input E of plane x+1 goes to sE1[PAR ^ 1], read next plane
  // after this plane's barriers)
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    L.sE1[PAR ^ 1][w][0][c * LPR + lane] = cur.e[0][c];
    L.sE1[PAR ^ 1][w][1][c * LPR + lane] = cur.e[2][c];
  }
  float eup[2][CPL];
  if (w < NW - 1) {
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      eup[0][c] = L.sE1[PAR][w + 1][0][c * LPR + lane];
      eup[1][c] = L.sE1[PAR][w + 1][1][c * LPR + lane];
    }
  } else {
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      eup[0][c] = cur.up[0][c];
      eup[1][c] = cur.up[1][c];
    }
  }
  float hn1[3][CPL], psH1[2][CPL];
  {
    float psH[2][CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      psH[0][c] = cur.ps[0][c];
      psH[1][c] = cur.ps[1][c];
    }
    h_row<CPL, LPR>(hn1, psH1, s.e0, cur.e, eup, cur.h, psH, zc, zb, Z, a.dt, act);
  }
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    L.sH1[w][0][c * LPR + lane] = hn1[0][c];
    L.sH1[w][1][c * LPR + lane] = hn1[2][c];
  }
  lds_barrier();

  // ---- B: level-1 E at plane x, then level-2 H at plane x-1
  float en1[3][CPL], psE1[2][CPL];
  if (e_rows) {
    float hdn[2][CPL], hp[2][CPL], psE[2][CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      hdn[0][c] = L.sH1[w - 1][0][c * LPR + lane];
      hdn[1][c] = L.sH1[w - 1][1][c * LPR + lane];
      hp[0][c] = s.h1m[1][c];
      hp[1][c] = s.h1m[2][c];
      psE[0][c] = cur.ps[2][c];
      psE[1][c] = cur.ps[3][c];
    }
    e_row<CPL, LPR>(en1, psE1, s.e0, hn1, hdn, hp, cur.cb, cur.sig, psE, zc, a, a.n, x, y, zb, lane,
               act);
    if ((a.flags & (SNAP1 | CAPTURE)) && own && act && x >= xa && x < xb)
      store_rare<CPL, vec>(a, a.flags & SNAP1, a.flags & CAPTURE, a.n, x, y, zb, en1, s.e0);
  } else {
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      en1[0][c] = en1[1][c] = en1[2][c] = 0.0f;
      psE1[0][c] = psE1[1][c] = 0.0f;
    }
  }
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    L.sE2[PAR][w][0][c * LPR + lane] = en1[0][c];
    L.sE2[PAR][w][1][c * LPR + lane] = en1[2][c];
  }
  float hn2[3][CPL], psH2[2][CPL];
  if (h2_rows && l2) {
    float e2up[2][CPL], psHm[2][CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      e2up[0][c] = L.sE2[PAR ^ 1][w + 1][0][c * LPR + lane];
      e2up[1][c] = L.sE2[PAR ^ 1][w + 1][1][c * LPR + lane];
      psHm[0][c] = s.psm[0][c];
      psHm[1][c] = s.psm[1][c];
    }
    h_row<CPL, LPR>(hn2, psH2, s.e1m, en1, e2up, s.h1m, psHm, zc, zb, Z, a.dt, row_valid);
  } else {
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      hn2[0][c] = hn2[1][c] = hn2[2][c] = 0.0f;
      psH2[0][c] = psH2[1][c] = 0.0f;
    }
  }
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    L.sH2[w][0][c * LPR + lane] = hn2[0][c];
    L.sH2[w][1][c * LPR + lane] = hn2[2][c];
  }
  lds_barrier();

  // ---- C: level-2 E at plane x-1 (the result on owned rows)
  const bool st = own && l2 && row_valid && x - 1 >= xa;
  float en2[3][CPL], psE2[2][CPL];
  if (own && l2 && row_valid) {
    float hdn[2][CPL], psEm[2][CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      hdn[0][c] = L.sH2[w - 1][0][c * LPR + lane];
      hdn[1][c] = L.sH2[w - 1][1][c * LPR + lane];
      psEm[0][c] = s.psm[2][c];
      psEm[1][c] = s.psm[3][c];
    }
    e_row<CPL, LPR>(en2, psE2, s.e1m, hn2, hdn, s.h2p, s.cbm, s.sigm, psEm, zc, a, a.n + 1, x - 1, y,
               zb, lane, true);
    if (st && (a.flags & (SNAP2 | CAPTURE)))
      store_rare<CPL, vec>(a, a.flags & SNAP2, a.flags & CAPTURE, a.n + 1, x - 1, y, zb, en2,
                           s.e1m);
  } else {
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      en2[0][c] = en2[1][c] = en2[2][c] = 0.0f;
      psE2[0][c] = psE2[1][c] = 0.0f;
    }
  }
  {
    const unsigned ob = st && lane_ok ? (unsigned)((((x - 1) * a.Y + y) * Z + zb) * 4) : kOff;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      bstore<CPL>(en2[k], O.e, ob, k * O.N4);
      bstore<CPL>(hn2[k], O.h, ob, k * O.N4);
    }
    const int pb = ((x - 1) * a.Y + y) * a.NP;
    if (CPL == 2 && a.pair) {
      const unsigned vp = st && zc.pml[0] ? (unsigned)((pb + zc.pidx[0]) * 4) : kOff;
      bstore<CPL>(psH2[0], O.psi, vp, 0);
      bstore<CPL>(psH2[1], O.psi, vp, O.P4);
      bstore<CPL>(psE2[0], O.psi, vp, 2 * O.P4);
      bstore<CPL>(psE2[1], O.psi, vp, 3 * O.P4);
    } else {
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        const unsigned vp = st && zc.pml[c] ? (unsigned)((pb + zc.pidx[c]) * 4) : kOff;
        bstore1(psH2[0][c], O.psi, vp, 0);
        bstore1(psH2[1][c], O.psi, vp, O.P4);
        bstore1(psE2[0][c], O.psi, vp, 2 * O.P4);
        bstore1(psE2[1][c], O.psi, vp, 3 * O.P4);
      }
    }
  }

  // ---- carry to plane x+1
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    s.h2p[0][c] = hn2[1][c];
    s.h2p[1][c] = hn2[2][c];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      s.e1m[k][c] = en1[k][c];
      s.h1m[k][c] = hn1[k][c];
      s.cbm[k][c] = cur.cb[k][c];
      s.e0[k][c] = cur.e[k][c];
    }
    s.psm[0][c] = psH1[0][c];
    s.psm[1][c] = psH1[1][c];
    s.psm[2][c] = psE1[0][c];
    s.psm[3][c] = psE1[1][c];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) s.sigm[k] = cur.sig[k];
}

// One tile (WY rows x LX planes) of one simulation for steps a.n, a.n+1.
template <int CPL, int WY, bool vec, int RPW = 1>
__device__ __forceinline__ void stream2_tile(const Stream2Args& a, int tile,
                                             Lds2<CPL, WY, RPW>& L) {
  constexpr int NW = WY + 3;
  constexpr int LPR = 64 / RPW;
  const int Z = a.Z;
  // Row w of the tile and lane within the row: one row per wave (the row is
  // wave-uniform), or two rows per wave, one per 32-lane half.
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = RPW == 1 ? (int)(threadIdx.x & 63) : (int)(threadIdx.x & (LPR - 1));
  const int w = RPW == 1 ? wave : wave * RPW + (int)((threadIdx.x & 63) / LPR);
  const int zb = lane * CPL;
  const int ty = tile % a.nty, tx = tile / a.nty;
  const int y = ty * WY - 2 + w;
  const bool row_valid = y >= 0 && y < a.Y;
  const int yc = min(max(y, 0), a.Y - 1);  // address-safe row
  const int xa = tx * a.LX;
  const int xb = min(xa + a.LX, a.X);

  const ZCoef<CPL> zc = make_zcoef<CPL>(a, zb);

  Carry2<CPL> s;
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
#pragma unroll
    for (int k = 0; k < 3; ++k) s.e1m[k][c] = s.h1m[k][c] = s.cbm[k][c] = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) s.psm[k][c] = 0.0f;
    s.h2p[0][c] = s.h2p[1][c] = 0.0f;
  }
  s.sigm[0] = s.sigm[1] = s.sigm[2] = 0.0f;
  const int xs = max(xa - 2, 0);
  {
    const int o = (xs * a.Y + yc) * Z;
#pragma unroll
    for (int k = 0; k < 3; ++k) load_col<CPL, vec>(s.e0[k], a.Ein + (k * a.N + o), zb, Z, row_valid);
  }
  Plane2<CPL> cur, nxt;
  fetch2<CPL, vec, RPW>(cur, a, xs, yc, w == NW - 1, zb, row_valid, w >= 1, zc);
  // Input Ex, Ez of the first plane for the row below's level-1 H.
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    L.sE1[0][w][0][c * LPR + lane] = s.e0[0][c];
    L.sE1[0][w][1][c * LPR + lane] = s.e0[2][c];
  }
  lds_barrier();
  const Out2 O = make_out2(a);
  const bool lane_ok = zb < Z;
  int par = 0;
  for (int x = xs; x <= xb; ++x) {
    plane2<CPL, WY, vec, RPW>(a, O, L, s, cur, nxt, zc, x, xa, xb, w, y, yc, lane, zb, row_valid,
                         lane_ok, par);
    cur = nxt;
    par ^= 1;
  }
}

template <int CPL, int WY, bool vec, int RPW = 1>
__global__ __launch_bounds__(64 * (WY + 3) / RPW) void k_stream2(Stream2Args a) {
  __shared__ Lds2<CPL, WY, RPW> L;
  // Consecutive workgroup ids go round-robin over the 8 XCDs; remap so that
  // tiles adjacent in y (which share ghost rows) run on the same XCD.
  int tile = blockIdx.x;
  if ((a.ntiles & 7) == 0) tile = (tile & 7) * (a.ntiles >> 3) + (tile >> 3);
  stream2_tile<CPL, WY, vec, RPW>(a, tile, L);
}

__device__ inline int uniform_load_int(const int* p) {
  return *(const __attribute__((address_space(4))) int*)p;
}

// Batched two-step kernel: persistent workgroups take (simulation, tile) work
// items from a device counter until the launch's items are exhausted; the
// simulation of item t is the last s with tile_start[s] <= t (tile counts are
// prefix-summed over simulations of uneven size).  args holds two views per
// simulation (which buffer set is old), selected by the parity of the step
// pair.  Every workgroup exits once the counter passes `total`.
template <int CPL, int WY, bool vec>
__global__ __launch_bounds__(64 * (WY + 3)) void k_stream2_batch(
    const Stream2Args* __restrict__ args, const int* __restrict__ tile_start, int nsims,
    int total, int* __restrict__ counter, const int* __restrict__ step_base, int step_off) {
  __shared__ Lds2<CPL, WY> L;
  __shared__ int s_item, s_next;
  if (threadIdx.x == 0) s_next = 0;
  // First step of the pair: step_off, plus a device-held base when the launch
  // is a node of a replayed graph chunk.
  const int step = (step_base ? uniform_load_int(step_base) : 0) + step_off;
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
    Stream2Args a = args[2 * lo + ((step >> 1) & 1)];
    if (step >= a.nsteps) continue;
    a.n = step;
    a.flags = (out_slot(a.rare, step) >= 0 ? SNAP1 : 0) |
              (out_slot(a.rare, step + 1) >= 0 ? SNAP2 : 0) | (a.rare->cap ? CAPTURE : 0);
    stream2_tile<CPL, WY, vec>(a, t - tile_start[lo], L);
  }
}

template <int CPL, int WY, bool vec, int RPW = 1>
static void launch2(Stream2Args& a, hipStream_t s) {
  static_assert((WY + 3) % RPW == 0, "rows of a tile must fill whole waves");
  hipLaunchKernelGGL((k_stream2<CPL, WY, vec, RPW>), dim3(a.ntiles), dim3(64 * (WY + 3) / RPW),
                     0, s, a);
}

// Whether the two-step kernel handles this problem (vector loads needed).
// Its store resources address at most 2^31 bytes: the three components of E
// (or of H) of one level, 12 B a cell, must fit, i.e. up to about 178M cells
// (the 32-bit element offsets of its loads reach 6 N, within 2^31 there), and
// so must the four psi components of one set (16 B per PML cell).
bool fits_rsrc(const Geom& g) {
  return g.N * 12 + 1024 <= (1LL << 31) && 16LL * g.X * g.Y * g.NP + 1024 <= (1LL << 31);
}

bool stream2_supported(const Geom& g) {
  return fits_rsrc(g) && g.Z <= 256 &&
         ((g.Z <= 64) || (g.Z <= 96 && g.Z % 3 == 0) || (g.Z <= 128 && g.Z % 2 == 0) ||
          (g.Z % 4 == 0));
}

// Heights 64 < zz <= 96 that are multiples of 3 run three cells per lane
// with two rows per wave (32 lanes each): every lane holds cells, where two
// cells per lane leave some lanes empty (a quarter at 96).  Where the
// LDS-ring kernel takes the height (multiples of 4, at most 32 PML cells;
// pjz's default budget, 128 - 32 = 96, is one) it is faster even so: 107.6
// against 100.7 Gcell-updates/s at 512x512x96.  tile_y 4, 8, 12 or 13 asks
// for the two-cells-per-lane kernels (even heights).
static bool use_cpl3(const Geom& g, const pjz_fdtd_problem& p) {
  if (g.Z <= 64 || g.Z > 96 || g.Z % 3 != 0) return false;
  if (stream2_ring_supported(g)) return false;
  const bool cpl2_asked = p.tile_y == 4 || p.tile_y == 8 || p.tile_y == 12 || p.tile_y == 13;
  return !(cpl2_asked && g.Z % 2 == 0);
}

// Runs steps [r.begin, r.end) two at a time; an odd last step uses the
// one-step kernel (same buffer sets).
int device_cus() {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  return cus;
}

// Kernel arguments of one simulation with buffer set `cur` as the old level.
Stream2Rare rare_of(const pjz_fdtd_problem& p, const OutWin& win, float* capture,
                    int cap_begin) {
  Stream2Rare r;
  r.out = win.out;
  r.cap = capture;
  r.cap_begin = cap_begin;
  r.wx0 = win.x0;
  r.wy0 = win.y0;
  r.wz0 = win.z0;
  r.wex = win.ex;
  r.wey = win.ey;
  r.wez = win.ez;
  r.ostart = p.out_start;
  r.ostop = p.out_stop;
  r.ostep = p.out_step;
  return r;
}

__global__ void k_put_rare(Stream2Rare* dst, Stream2Rare v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = v;
}

Stream2Args make_args2(const Geom& g, const pjz_fdtd_problem& p, const float* mask,
                       const Src& src, const OutWin& win, char* ws, const Layout& L, int WY,
                       int LX, int cur) {
  Stream2Args a;
  float* psi = (float*)(ws + L.psi);
  const int P = g.X * g.Y * g.NP;
  float* E[2] = {(float*)(ws + L.E0), (float*)(ws + L.E1)};
  a.Ein = E[cur];
  a.Eout = E[1 - cur];
  a.psiIn = psi + (cur ? PSI_SET : 0) * (long long)P;
  a.psiOut = psi + (cur ? 0 : PSI_SET) * (long long)P;
  a.cb = (const float*)(ws + L.cb);
  a.zt = (const float*)(ws + L.zt);
  a.mask = mask;
  a.sfield = src.field;
  a.swave = src.wave;
  a.saxis = src.axis;
  a.spos = src.pos;
  a.squad = src.quad;
  a.rare = nullptr;
  a.X = g.X;
  a.Y = g.Y;
  a.Z = g.Z;
  a.lo = g.lo;
  a.hi = g.hi;
  a.NP = g.NP;
  a.N = (int)g.N;
  a.P = P;
  a.hoff = (int)((L.H0 - L.E0) / sizeof(float));
  a.dt = p.dt;
  a.n = 0;
  a.flags = 0;
  a.pair = g.Z > 64 && g.Z <= 128 && g.Z % 2 == 0 && g.lo % 2 == 0 && g.hi % 2 == 0;
  a.LX = LX;
  a.nty = (g.Y + WY - 1) / WY;
  a.ntiles = a.nty * ((g.X + LX - 1) / LX);
  a.nsteps = steps_of(p);
  return a;
}

hipError_t run_stream2(const Geom& g, const pjz_fdtd_problem& p, const float* mask,
                       const Src& src, const OutWin& win, char* ws, const Layout& L,
                       const StepRange& r, hipStream_t s) {
  if (!stream2_supported(g)) return hipErrorInvalidValue;
  const int cpl = g.Z <= 64 ? 1 : use_cpl3(g, p) ? 3 : (g.Z <= 128 ? 2 : 4);
  // Row tiles with a compiled kernel: 8 (4 at CPL = 4); CPL = 2 also 4, 12,
  // 13; CPL = 1 also 13 (16 waves, which its register use lets share a CU).
  int WY = cpl == 4 ? 4 : cpl == 3 ? 13 : 8;
  if (cpl == 2 && (p.tile_y == 4 || p.tile_y == 12 || p.tile_y == 13)) WY = p.tile_y;
  // One cell per lane: 13-row tiles (16 waves) unless they cannot give every
  // CU one even at 8-plane chunks (measured at 512^2, 300^2 and 200^2 x 64:
  // 17-29% faster than 8 rows).
  if (cpl == 1 && p.tile_y == 0 &&
      (long long)((g.Y + 12) / 13) * ((g.X + 7) / 8) >= device_cus())
    WY = 13;
  if (cpl == 1 && p.tile_y == 13) WY = 13;
  int LX = p.tile_x;
  if (LX <= 0) {
    // x chunks sized so the grid fills the CUs in whole rounds.
    LX = grid_chunk_x(g.X, (g.Y + WY - 1) / WY, device_cus(), 2);
  }
  Stream2Args a = make_args2(g, p, mask, src, win, ws, L, WY, LX, 0);
  // 8-row tiles at two cells per lane take their inputs through an LDS ring
  // two planes deep where the height and PML widths allow.
  const bool ring = cpl == 2 && WY == 8 && stream2_ring_supported(g);
  Stream2Rare* rare = (Stream2Rare*)(ws + L.rare);
  hipLaunchKernelGGL(k_put_rare, dim3(1), dim3(1), 0, s, rare,
                     rare_of(p, win, r.capture, r.begin));
  a.rare = rare;
  float* psi = (float*)(ws + L.psi);
  const int P = g.X * g.Y * g.NP;
  float* E[2] = {(float*)(ws + L.E0), (float*)(ws + L.E1)};
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
    a.Eout = E[1 - c];
    a.psiIn = psi + (c ? PSI_SET : 0) * (long long)P;
    a.psiOut = psi + (c ? 0 : PSI_SET) * (long long)P;
    a.n = n;
    a.flags = (slot(n) >= 0 ? SNAP1 : 0) | (slot(n + 1) >= 0 ? SNAP2 : 0) |
              (r.capture ? CAPTURE : 0);
    if (cpl == 3) launch2<3, 13, true, 2>(a, s);
    else if (cpl == 1 && WY == 13) launch2<1, 13, true>(a, s);
    else if (cpl == 1) launch2<1, 8, true>(a, s);
    else if (cpl == 4) launch2<4, 4, true>(a, s);
    else if (WY == 4) launch2<2, 4, true>(a, s);
    else if (WY == 12) launch2<2, 12, true>(a, s);
    else if (WY == 13) launch2<2, 13, true>(a, s);
    else if (ring) (void)launch_stream2_ring(a, s);
    else launch2<2, 8, true>(a, s);
    *r.cur = 1 - c;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || n >= r.end) return e;
  StepRange tail{n, r.end, r.cur, r.capture ? r.capture + (long long)(n - r.begin) * f3 : nullptr};
  return run_stream(g, p, mask, src, win, ws, L, tail, s);
}

// ---------------------------------------------------------------- batches
// Simulations of a batch are grouped by cells per lane (zz <= 64: 1, even
// zz <= 128: 2); each group runs as one persistent launch per step pair.
bool batch2_supported(const Geom& g) {
  return fits_rsrc(g) && (g.Z <= 64 || (g.Z <= 128 && g.Z % 2 == 0));
}

// Rows per batch tile by cells-per-lane class: zz <= 64 (CPL = 1) takes 13
// rows (16 waves at ~120 registers, one workgroup filling a CU), zz <= 128
// takes 8 (11 waves at 168 registers).
static constexpr int kBatchWY1 = 13;       // cells per lane 1: rows per tile
static constexpr int kBatchWY1Small = 8;   // ... when too few tiles fill the device
static constexpr int kBatchWY2 = 8;
static constexpr int kBatchLX = 128;

// Workgroups of a persistent kernel that fit on one CU at once (occupancy
// query, once per kernel): the persistent grid is that many per CU.
template <typename K>
static int resident_per_cu(K kernel, int threads) {
  int b = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, reinterpret_cast<const void*>(kernel),
                                                   threads, 0) != hipSuccess || b < 1)
    b = 1;
  return b;
}

template <int CPL, int WY>
static int batch2_resident() {
  static const int n = resident_per_cu(&k_stream2_batch<CPL, WY, true>, 64 * (WY + 3));
  return n;
}

// x chunk length of a batch's tiles.  The persistent workgroups take tiles
// from the queue in order (simulation by simulation, tile t of a simulation
// at row tile t % nty, x chunk t / nty), each taking about min(LX, rest of
// X) + 2 planes of time (two ghost planes).  Simulate that greedy schedule
// on `slots` workgroups for each candidate LX and keep the longest chunk
// within 3% of the shortest makespan (longer chunks re-read fewer ghost
// planes, which the model does not charge).
static long long batch_makespan(const int* xs, const int* nty, int n, int slots, int lx,
                                int ghost) {
  std::priority_queue<long long, std::vector<long long>, std::greater<long long>> free_at;
  for (int w = 0; w < slots; ++w) free_at.push(0);
  long long span = 0;
  for (int i = 0; i < n; ++i) {
    const int ntx = (xs[i] + lx - 1) / lx;
    for (int tx = 0; tx < ntx; ++tx) {
      const long long d = std::min(lx, xs[i] - tx * lx) + ghost;
      for (int ty = 0; ty < nty[i]; ++ty) {
        const long long t = free_at.top() + d;
        free_at.pop();
        free_at.push(t);
        span = std::max(span, t);
      }
    }
  }
  return span;
}

int batch_chunk_x(const int* xs, const int* nty, int n, int slots, int max_lx, int ghost) {
  // Down to 8 planes, the single-simulation minimum: a batch of one or a few
  // small simulations needs short chunks to put a tile on every workgroup.
  static const int cands[] = {128, 96, 64, 48, 32, 24, 16, 12, 8};
  std::vector<long long> cost;
  std::vector<int> lxs;
  for (int lx : cands) {
    if (lx > max_lx) continue;
    cost.push_back(batch_makespan(xs, nty, n, slots, lx, ghost));
    lxs.push_back(lx);
    // Eight or more rounds of tiles already keep every workgroup busy to
    // within an eighth of a round: shorter chunks only add ghost planes.
    long long tiles = 0;
    for (int i = 0; i < n; ++i) tiles += (long long)nty[i] * ((xs[i] + lx - 1) / lx);
    if (tiles >= 8LL * slots) break;
  }
  if (lxs.empty()) return max_lx;
  const long long best = *std::min_element(cost.begin(), cost.end());
  for (size_t c = 0; c < lxs.size(); ++c)  // candidates run longest first
    if (cost[c] * 100 <= best * 103) return lxs[c];
  return lxs[0];
}
// x chunk length of a single simulation's grid (nty row tiles per chunk,
// `ghost` recomputed planes per chunk).  The device runs `slots` workgroups
// at a time, so a grid just above a multiple of `slots` ends with a nearly
// empty extra round of tiles.  A chunk length costs rounds x (planes + ghost
// planes + 2 for a tile's start-up); keep the longest chunk within 3% of the
// cheapest.  Chunks are at least 8 planes (or all of X).
int grid_chunk_x(int X, int nty, int slots, int ghost) {
  slots = std::max(slots, 1);
  std::vector<std::pair<long long, int>> c;  // (cost, chunk), longest first
  for (int ntx = 1;; ++ntx) {
    const int lx = (X + ntx - 1) / ntx;
    if (ntx > 1 && lx < 8) break;
    if (c.empty() || c.back().second != lx) {
      const long long tiles = (long long)nty * ((X + lx - 1) / lx);
      c.push_back({(tiles + slots - 1) / slots * (long long)(lx + ghost + 2), lx});
    }
    if (lx <= 8) break;
  }
  long long best = c.front().first;
  for (const auto& e : c) best = std::min(best, e.first);
  for (const auto& e : c)
    if (e.first * 100 <= best * 103) return e.second;
  return c.front().second;
}

static constexpr int kMaxChunkPairs = 64;  // step pairs per captured graph chunk

size_t batch2_meta_bytes(const BatchItem* items, int n) {
  int npairs = 0;
  for (int i = 0; i < n; ++i) npairs = std::max(npairs, (steps_of(*items[i].p) + 1) / 2);
  return align256(sizeof(Stream2Args) * 2 * (size_t)n) + align256(sizeof(Stream2Rare) * (size_t)n) +
         align256(sizeof(int) * (size_t)(n + 2) * 2) +
         align256(sizeof(int) * 2 * kQueues * (size_t)std::max(npairs, 1)) + align256(sizeof(int)) +
         align256(sizeof(int) * 2 * kQueues * (size_t)kMaxChunkPairs);
}

hipError_t batch2_prepare(const BatchItem* items, int n, char* meta, Batch2Plan& plan,
                          hipStream_t s, bool forward_only) {
  std::vector<int> cls[2];
  plan.n = n;
  plan.npairs = 0;
  plan.depth0 = 2;
  plan.nlaunch0 = 0;
  plan.meta = meta;
  plan.cus = device_cus();
  plan.graph = nullptr;
  plan.exec = nullptr;
  plan.chunk = std::min(std::max(items[0].p->graph_steps / 2, 0), kMaxChunkPairs);
  plan.ring1 = true;
  for (int i = 0; i < n; ++i) {
    if (!batch2_supported(items[i].g)) return hipErrorInvalidValue;
    cls[items[i].g.Z <= 64 ? 0 : 1].push_back(i);
    if (items[i].g.Z > 64) plan.ring1 = plan.ring1 && stream2_ring_supported(items[i].g);
    plan.npairs = std::max(plan.npairs, (steps_of(*items[i].p) + 1) / 2);
  }
  // Host image of the argument views, launch data and tile prefix sums,
  // copied once.  Layout: args (2 per sim, class order) | rare (class order)
  // | tile starts (n_c + 1 per class) | queue counters (2 x kQueues per
  // pair) | graph chunk step base | graph chunk queue counters.
  plan.args_bytes = align256(sizeof(Stream2Args) * 2 * (size_t)n);
  plan.rare_bytes = align256(sizeof(Stream2Rare) * (size_t)n);
  plan.ts_bytes = align256(sizeof(int) * (size_t)(n + 2) * 2);
  std::vector<char> host(plan.args_bytes + plan.rare_bytes + plan.ts_bytes, 0);
  Stream2Args* hargs = (Stream2Args*)host.data();
  Stream2Rare* hrare = (Stream2Rare*)(host.data() + plan.args_bytes);
  Stream2Rare* drare = (Stream2Rare*)(meta + plan.args_bytes);
  int* hts = (int*)(host.data() + plan.args_bytes + plan.rare_bytes);
  int off = 0;
  for (int c = 0; c < 2; ++c) {
    plan.base[c] = off;
    plan.count[c] = (int)cls[c].size();
    // Row tiles: 8 at two cells per lane; at one cell per lane 13 (16 waves,
    // fewer ghost rows per owned row) unless 13-row tiles cannot give every
    // workgroup one even at the shortest x chunk (8 planes), where 8-row
    // tiles (11 waves) finish sooner.  Measured on 128x128x64 simulations:
    // one alone runs 11% faster with 8-row tiles, two or more 3-26% slower.
    std::vector<int> xs;
    for (int i : cls[c]) xs.push_back(items[i].g.X);
    auto chunk = [&](int wy, int resident, int ghost) {
      std::vector<int> ntys;
      for (int i : cls[c]) ntys.push_back((items[i].g.Y + wy - 1) / wy);
      return xs.empty() ? kBatchLX
                        : batch_chunk_x(xs.data(), ntys.data(), (int)xs.size(),
                                        plan.cus * resident, kBatchLX, ghost);
    };
    int WY = c == 0 ? kBatchWY1 : kBatchWY2;
    if (c == 0) {
      // The caller may fix the rows per tile: tile_y = 8 or 13 on every
      // simulation of the class.
      int forced = 0;
      for (int i : cls[c]) {
        const int ty = items[i].p->tile_y;
        const bool ok = ty == kBatchWY1 || ty == kBatchWY1Small;
        forced = ok && (forced == 0 || forced == ty) ? ty : -1;
        if (forced < 0) break;
      }
      long long tiles13 = 0;
      for (int i : cls[c])
        tiles13 += (long long)((items[i].g.Y + kBatchWY1 - 1) / kBatchWY1) *
                   ((items[i].g.X + 7) / 8);
      const bool small = forced > 0 ? forced == kBatchWY1Small
                                    : tiles13 < (long long)plan.cus * batch2_resident<1, kBatchWY1>();
      if (small) WY = kBatchWY1Small;
      // Forward-only batches without graph chunks take three steps per
      // launch at one cell per lane when the class's 11-row tiles fill every
      // workgroup at 8-plane chunks (as single calls do on grids that fill
      // the device).
      if (forward_only && plan.chunk == 0 && forced <= 0 && !small) {
        const int wy3 = deep_batch_rows();
        long long tiles = 0;
        int maxsteps = 0;
        for (int i : cls[c]) {
          tiles += (long long)((items[i].g.Y + wy3 - 1) / wy3) * ((items[i].g.X + 7) / 8);
          maxsteps = std::max(maxsteps, steps_of(*items[i].p));
        }
        if (tiles >= (long long)plan.cus * deep_batch_resident()) {
          WY = wy3;
          plan.depth0 = 3;
          plan.nlaunch0 = (maxsteps + 2) / 3;
        }
      }
    }
    const int cLX = c == 1 ? chunk(kBatchWY2,
                                   plan.ring1 ? stream2_ring_batch_resident()
                                              : batch2_resident<2, kBatchWY2>(),
                                   2)
                    : plan.depth0 == 3 ? chunk(WY, deep_batch_resident(), 3)
                    : WY == kBatchWY1Small ? chunk(WY, batch2_resident<1, kBatchWY1Small>(), 2)
                                           : chunk(WY, batch2_resident<1, kBatchWY1>(), 2);
    plan.wy[c] = WY;
    int t = 0;
    for (size_t k = 0; k < cls[c].size(); ++k) {
      const BatchItem& it = items[cls[c][k]];
      const int LX = std::min(cLX, it.g.X);
      hrare[off + k] = rare_of(*it.p, it.win, it.capture, 0);
      for (int cur = 0; cur < 2; ++cur) {
        hargs[2 * (off + k) + cur] =
            make_args2(it.g, *it.p, it.mask, it.src, it.win, it.ws, *it.L, WY, LX, cur);
        hargs[2 * (off + k) + cur].rare = drare + off + k;
      }
      hts[off + c + k] = t;
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

// The launches of one step pair (one per class), queue counters
// ctr[c * kQueues ..] of class c (zero on entry), first step *step_base + step_off (step_base may be null).
static void launch_pair(const Batch2Plan& plan, hipStream_t s, int* ctr, const int* step_base,
                        int step_off, int classes = 3) {
  const Stream2Args* dargs = (const Stream2Args*)plan.meta;
  const int* dts = (const int*)(plan.meta + plan.args_bytes + plan.rare_bytes);
  for (int c = 0; c < 2; ++c) {
    const int nc = plan.count[c];
    if (nc == 0 || !(classes & (1 << c))) continue;
    const Stream2Args* a = dargs + 2 * plan.base[c];
    const int* ts = dts + plan.base[c] + c;
    if (c == 0 && plan.wy[0] == kBatchWY1Small) {
      const int grid = std::min(plan.total[c], plan.cus * batch2_resident<1, kBatchWY1Small>());
      hipLaunchKernelGGL((k_stream2_batch<1, kBatchWY1Small, true>), dim3(grid),
                         dim3(64 * (kBatchWY1Small + 3)), 0, s, a, ts, nc, plan.total[c],
                         ctr + c * kQueues, step_base, step_off);
    } else if (c == 0) {
      const int grid = std::min(plan.total[c], plan.cus * batch2_resident<1, kBatchWY1>());
      hipLaunchKernelGGL((k_stream2_batch<1, kBatchWY1, true>), dim3(grid),
                         dim3(64 * (kBatchWY1 + 3)), 0, s, a, ts, nc, plan.total[c], ctr + c * kQueues,
                         step_base, step_off);
    } else if (plan.ring1) {
      const int grid = std::min(plan.total[c], plan.cus * stream2_ring_batch_resident());
      launch_stream2_ring_batch(a, ts, nc, plan.total[c], ctr + c * kQueues, step_base, step_off,
                                grid, s);
    } else {
      const int grid = std::min(plan.total[c], plan.cus * batch2_resident<2, kBatchWY2>());
      hipLaunchKernelGGL((k_stream2_batch<2, kBatchWY2, true>), dim3(grid),
                         dim3(64 * (kBatchWY2 + 3)), 0, s, a, ts, nc, plan.total[c], ctr + c * kQueues,
                         step_base, step_off);
    }
  }
}

__global__ void k_set_int(int* p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

__global__ void k_add_int(int* p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p += v;
}

static int* chunk_base(const Batch2Plan& plan) {
  int* counters = (int*)(plan.meta + plan.args_bytes + plan.rare_bytes + plan.ts_bytes);
  return (int*)((char*)counters +
                align256(sizeof(int) * 2 * kQueues * (size_t)std::max(plan.npairs, 1)));
}

// Captures one chunk of plan.chunk step pairs as a graph: clear the chunk's
// queue counters (a fill kernel: no memset nodes), the pair launches reading
// their first step from the device-held base, then advance the base, so the
// same executable graph replays for every chunk.
static hipError_t build_chunk_graph(Batch2Plan& plan) {
  int* base = chunk_base(plan);
  int* ctr = base + 256 / sizeof(int);
  hipStream_t cs;
  hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if (e != hipSuccess) return e;
  e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    (void)launch_zero((float*)ctr, 2 * kQueues * (size_t)plan.chunk, cs);
    for (int i = 0; i < plan.chunk; ++i) launch_pair(plan, cs, ctr + 2 * kQueues * i, base, 2 * i);
    hipLaunchKernelGGL(k_add_int, dim3(1), dim3(1), 0, cs, base, 2 * plan.chunk);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(cs, &g);
    if (e == hipSuccess) {
      plan.graph = g;
      e = hipGraphInstantiate(&plan.exec, g, nullptr, nullptr, 0);
    }
  }
  (void)hipStreamDestroy(cs);
  return e;
}

hipError_t batch2_pairs(Batch2Plan& plan, int m0, int m1, hipStream_t s) {
  m1 = std::min(m1, plan.npairs);
  if (m0 >= m1) return hipSuccess;
  hipError_t e = hipSuccess;
  if (plan.chunk > 0 && m1 - m0 >= plan.chunk) {
    if (!plan.exec) e = build_chunk_graph(plan);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, chunk_base(plan), 2 * m0);
    for (; m1 - m0 >= plan.chunk && e == hipSuccess; m0 += plan.chunk)
      e = hipGraphLaunch(plan.exec, s);
    if (e != hipSuccess || m0 >= m1) return e;
  }
  int* counters = (int*)(plan.meta + plan.args_bytes + plan.rare_bytes + plan.ts_bytes);
  // Queue counters of these pairs start at zero (a fill kernel).
  e = launch_zero((float*)(counters + 2 * kQueues * m0), 2 * kQueues * (size_t)(m1 - m0), s);
  if (e != hipSuccess) return e;
  for (int m = m0; m < m1; ++m) launch_pair(plan, s, counters + 2 * kQueues * m, nullptr, 2 * m);
  return hipGetLastError();
}

hipError_t batch2_forward(Batch2Plan& plan, hipStream_t s) {
  if (plan.depth0 != 3) return batch2_pairs(plan, 0, plan.npairs, s);
  // zz <= 64 class: three steps a launch; the other class in pairs.  Launch
  // t of the first class uses the queue counters of pair slot t, class 0;
  // pair m of the second those of slot m, class 1 (nlaunch0 <= npairs).
  int* counters = (int*)(plan.meta + plan.args_bytes + plan.rare_bytes + plan.ts_bytes);
  hipError_t e = launch_zero((float*)counters, 2 * kQueues * (size_t)plan.npairs, s);
  if (e != hipSuccess) return e;
  const Stream2Args* a = (const Stream2Args*)plan.meta + 2 * plan.base[0];
  const int* ts = (const int*)(plan.meta + plan.args_bytes + plan.rare_bytes) + plan.base[0];
  const int grid = std::min(plan.total[0], plan.cus * deep_batch_resident());
  for (int t = 0; t < plan.nlaunch0; ++t)
    launch_deep_batch(a, ts, plan.count[0], plan.total[0], counters + 2 * kQueues * t, 3 * t,
                      grid, s);
  for (int m = 0; m < plan.npairs; ++m)
    launch_pair(plan, s, counters + 2 * kQueues * m, nullptr, 2 * m, 2);
  return hipGetLastError();
}

hipError_t batch2_release(Batch2Plan& plan, hipStream_t s) {
  hipError_t e = hipSuccess;
  if (plan.exec) {
    // The graph may still be running: destroy it once the stream drained.
    e = hipStreamSynchronize(s);
    (void)hipGraphExecDestroy(plan.exec);
    plan.exec = nullptr;
  }
  if (plan.graph) {
    (void)hipGraphDestroy(plan.graph);
    plan.graph = nullptr;
  }
  return e;
}

__global__ void k_capture_from(Stream2Rare* r, int n, int begin) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i].cap_begin = begin;
}

hipError_t batch2_capture_from(const Batch2Plan& plan, int begin, hipStream_t s) {
  Stream2Rare* r = (Stream2Rare*)(plan.meta + plan.args_bytes);
  hipLaunchKernelGGL(k_capture_from, dim3((plan.n + 255) / 256), dim3(256), 0, s, r, plan.n,
                     begin);
  return hipGetLastError();
}

}  // namespace pjz
