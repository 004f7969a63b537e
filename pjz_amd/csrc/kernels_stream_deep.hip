// Deeper temporal blocking: one launch advances T time steps (T = 3 or 4)
// while reading and writing each field once.  The two-step kernel
// (kernels_stream2.hip) is the T = 2 case of the same schedule; this file
// generalises it over the depth, with every level's state in registers.
//
// Per workgroup: rows y0-T .. y0+WY+T-2 (NW = WY + 2T - 1 wavefronts, one y
// row each, whole z column per wave).  Sweeping x, level l (step n + l - 1)
// works on plane x_l = x - l + 1.  A plane of the sweep runs T + 1 phases
// separated by T LDS-only barriers:
//   phase 0:  level-1 H(x)           needs input E at x, x+1 and row y+1
//                                    (published one plane early)
//   phase p:  level-p E(x_p)         needs level-p H of row y-1 (phase p-1)
//   (1..T)    level-(p+1) H(x_p - 1) needs level-p E at x_p - 1 and x_p of
//                                    this row and of row y+1 at x_p - 1
//                                    (published one plane earlier)
// Level l's H runs on rows w = l-1 .. NW-l and its E on w = l .. NW-l (w the
// row within the tile), so level T's E covers exactly the owned rows; planes
// xa-T .. xa-1 are recomputed ghost planes.  Only level T's fields (and
// snapshots / captures of every level on owned cells) are written.
//
// Registers: a level carries its E, H, psi, cb and absorber of the previous
// plane, so registers grow with T.  At one cell per lane (zz <= 64) depth 3
// fits a 16-wave workgroup (WY = 11, ~111 registers) and depth 4 a 13-wave
// one (WY = 6).  At two cells per lane a third level does not fit beside the
// second without spilling (measured 3.4 x slower than two steps per launch
// at 512x512x128), so the deeper kernels are for zz <= 64.
#include "stream2_parts.h"

namespace pjz {

// State carried from plane x-1 to plane x; entry [l-1] belongs to level l.
template <int T, int CPL>
struct CarryT {
  float e0[3][CPL];         // input E at x
  float em[T][3][CPL];      // level-l E at x - l (l < T)
  float hm[T][3][CPL];      // level-l H at x - l (level T: Hy, Hz only)
  float cbm[T][3][CPL];     // cb at x - l (l < T)
  float psm[T][4][CPL];     // level-l psi (HX, HY, EX, EY) at x - l (l < T)
  float sigm[T][3];         // absorber at x - l (l < T)
};

// Row exchange buffers (cell c of lane l at [c * 64 + l]).  sE[l-1][parity]
// holds level-(l-1) Ex, Ez of each row (the input E for l = 1), read by
// level-l H of the row below; sH[l-1] holds level-l Hx, Hz, read by level-l E
// of the row above.
template <int T, int CPL, int WY>
struct LdsT {
  static constexpr int NW = WY + 2 * T - 1;
  float sE[T][2][NW][2][64 * CPL];
  float sH[T][NW][2][64 * CPL];
};

__host__ __device__ constexpr int snap_flag(int level) {
  return level == 1 ? SNAP1 : level == 2 ? SNAP2 : level == 3 ? SNAP3 : SNAP4;
}

template <int T, int CPL, int WY, bool vec>
__device__ __forceinline__ void planeT(const Stream2Args& a, const Out2& O, LdsT<T, CPL, WY>& L,
                                       CarryT<T, CPL>& s, const Plane2<CPL>& cur,
                                       Plane2<CPL>& nxt, const ZCoef<CPL>& zc, int x, int xs,
                                       int xa, int xb, int xe, int w, int y, int yc, int lane,
                                       int zb, bool row_valid, bool lane_ok, int PAR) {
  constexpr int NW = WY + 2 * T - 1;
  const int Z = a.Z;
  if (x + 1 <= xe)
    fetch2<CPL, vec>(nxt, a, x + 1, yc, w == NW - 1, zb, row_valid && x + 1 < a.X, w >= 1, zc);
  // Level l is computed where its plane x_l = x - l + 1 is in the domain and,
  // after a ghost start (xs > 0), far enough from it for its H to be exact
  // (x_l >= xs + l - 1).  From the domain's first plane every level is exact
  // (nothing lies below plane 0).
  bool lv[T];
#pragma unroll
  for (int l = 1; l <= T; ++l) {
    const int xl = x - l + 1;
    lv[l - 1] = xl >= (xs > 0 ? xs + l - 1 : 0) && xl < a.X;
  }

  float hn[T][3][CPL], psH[T][2][CPL], en[T][3][CPL], psE[T][2][CPL];

  // ---- phase 0: level-1 H at plane x
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    L.sE[0][PAR ^ 1][w][0][c * 64 + lane] = cur.e[0][c];
    L.sE[0][PAR ^ 1][w][1][c * 64 + lane] = cur.e[2][c];
  }
  {
    float eup[2][CPL], ps[2][CPL];
    if (w < NW - 1) {
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        eup[0][c] = L.sE[0][PAR][w + 1][0][c * 64 + lane];
        eup[1][c] = L.sE[0][PAR][w + 1][1][c * 64 + lane];
      }
    } else {
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        eup[0][c] = cur.up[0][c];
        eup[1][c] = cur.up[1][c];
      }
    }
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      ps[0][c] = cur.ps[0][c];
      ps[1][c] = cur.ps[1][c];
    }
    h_row<CPL>(hn[0], psH[0], s.e0, cur.e, eup, cur.h, ps, zc, zb, Z, a.dt, row_valid && lv[0]);
  }
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    L.sH[0][w][0][c * 64 + lane] = hn[0][0][c];
    L.sH[0][w][1][c * 64 + lane] = hn[0][2][c];
  }
  lds_barrier();

  // ---- phases 1..T
#pragma unroll
  for (int p = 1; p <= T; ++p) {
    const int xp = x - p + 1;
    // level-p E at plane xp
    if (w >= p && w <= NW - p && lv[p - 1] && row_valid) {
      float hdn[2][CPL], hp[2][CPL], eo[3][CPL], cb[3][CPL], pe[2][CPL], sig[3];
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        hdn[0][c] = L.sH[p - 1][w - 1][0][c * 64 + lane];
        hdn[1][c] = L.sH[p - 1][w - 1][1][c * 64 + lane];
        hp[0][c] = s.hm[p - 1][1][c];
        hp[1][c] = s.hm[p - 1][2][c];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          eo[k][c] = p == 1 ? s.e0[k][c] : s.em[p > 1 ? p - 2 : 0][k][c];
          cb[k][c] = p == 1 ? cur.cb[k][c] : s.cbm[p > 1 ? p - 2 : 0][k][c];
        }
        pe[0][c] = p == 1 ? cur.ps[2][c] : s.psm[p > 1 ? p - 2 : 0][2][c];
        pe[1][c] = p == 1 ? cur.ps[3][c] : s.psm[p > 1 ? p - 2 : 0][3][c];
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) sig[k] = p == 1 ? cur.sig[k] : s.sigm[p > 1 ? p - 2 : 0][k];
      e_row<CPL>(en[p - 1], psE[p - 1], eo, hn[p - 1], hdn, hp, cb, sig, pe, zc, a, a.n + p - 1,
                 xp, y, zb, lane, true);
      if ((a.flags & (snap_flag(p) | CAPTURE)) && w >= T && w <= NW - T && xp >= xa && xp < xb)
        store_rare<CPL, vec>(a, a.flags & snap_flag(p), a.flags & CAPTURE, a.n + p - 1, xp, y, zb,
                             en[p - 1], eo);
    } else {
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        en[p - 1][0][c] = en[p - 1][1][c] = en[p - 1][2][c] = 0.0f;
        psE[p - 1][0][c] = psE[p - 1][1][c] = 0.0f;
      }
    }
    if (p < T) {
      // level-(p+1) H at plane xp - 1 (q = p; the clamp only keeps the
      // indexes of the unrolled, never-taken p = T branch in range)
      const int q = p < T ? p : T - 1;
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        L.sE[q][PAR][w][0][c * 64 + lane] = en[p - 1][0][c];
        L.sE[q][PAR][w][1][c * 64 + lane] = en[p - 1][2][c];
      }
      if (w >= p && w <= NW - p - 1 && lv[q] && row_valid) {
        float eup[2][CPL], ps[2][CPL];
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          eup[0][c] = L.sE[q][PAR ^ 1][w + 1][0][c * 64 + lane];
          eup[1][c] = L.sE[q][PAR ^ 1][w + 1][1][c * 64 + lane];
          ps[0][c] = s.psm[p - 1][0][c];
          ps[1][c] = s.psm[p - 1][1][c];
        }
        h_row<CPL>(hn[q], psH[q], s.em[p - 1], en[p - 1], eup, s.hm[p - 1], ps, zc, zb, Z, a.dt,
                   true);
      } else {
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          hn[q][0][c] = hn[q][1][c] = hn[q][2][c] = 0.0f;
          psH[q][0][c] = psH[q][1][c] = 0.0f;
        }
      }
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        L.sH[q][w][0][c * 64 + lane] = hn[q][0][c];
        L.sH[q][w][1][c * 64 + lane] = hn[q][2][c];
      }
      lds_barrier();
    }
  }

  // ---- level-T results on owned rows at plane x - T + 1
  {
    const int xT = x - T + 1;
    const bool st = w >= T && w <= NW - T && lv[T - 1] && row_valid && xT >= xa;
    const unsigned ob = st && lane_ok ? (unsigned)(((xT * a.Y + y) * Z + zb) * 4) : kOff;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      bstore<CPL>(en[T - 1][k], O.e, ob, k * O.N4);
      bstore<CPL>(hn[T - 1][k], O.h, ob, k * O.N4);
    }
    const int pb = (xT * a.Y + y) * a.NP;
    if (CPL == 2 && a.pair) {
      const unsigned vp = st && zc.pml[0] ? (unsigned)((pb + zc.pidx[0]) * 4) : kOff;
      bstore<CPL>(psH[T - 1][0], O.psi, vp, 0);
      bstore<CPL>(psH[T - 1][1], O.psi, vp, O.P4);
      bstore<CPL>(psE[T - 1][0], O.psi, vp, 2 * O.P4);
      bstore<CPL>(psE[T - 1][1], O.psi, vp, 3 * O.P4);
    } else {
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        const unsigned vp = st && zc.pml[c] ? (unsigned)((pb + zc.pidx[c]) * 4) : kOff;
        bstore1(psH[T - 1][0][c], O.psi, vp, 0);
        bstore1(psH[T - 1][1][c], O.psi, vp, O.P4);
        bstore1(psE[T - 1][0][c], O.psi, vp, 2 * O.P4);
        bstore1(psE[T - 1][1][c], O.psi, vp, 3 * O.P4);
      }
    }
  }

  // ---- carry to plane x+1 (deepest levels first: cb and the absorber
  // shift one level down)
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    s.hm[T - 1][1][c] = hn[T - 1][1][c];
    s.hm[T - 1][2][c] = hn[T - 1][2][c];
#pragma unroll
    for (int l = T - 1; l >= 1; --l) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        s.em[l - 1][k][c] = en[l - 1][k][c];
        s.hm[l - 1][k][c] = hn[l - 1][k][c];
        s.cbm[l - 1][k][c] = l == 1 ? cur.cb[k][c] : s.cbm[l > 1 ? l - 2 : 0][k][c];
      }
      s.psm[l - 1][0][c] = psH[l - 1][0][c];
      s.psm[l - 1][1][c] = psH[l - 1][1][c];
      s.psm[l - 1][2][c] = psE[l - 1][0][c];
      s.psm[l - 1][3][c] = psE[l - 1][1][c];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) s.e0[k][c] = cur.e[k][c];
  }
#pragma unroll
  for (int l = T - 1; l >= 1; --l)
#pragma unroll
    for (int k = 0; k < 3; ++k) s.sigm[l - 1][k] = l == 1 ? cur.sig[k] : s.sigm[l > 1 ? l - 2 : 0][k];
}

// One tile (WY rows x LX planes) of one simulation for steps a.n .. a.n+T-1.
template <int T, int CPL, int WY, bool vec>
__device__ __forceinline__ void deep_tile(const Stream2Args& a, int tile, LdsT<T, CPL, WY>& L) {
  constexpr int NW = WY + 2 * T - 1;
  const int Z = a.Z;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int zb = lane * CPL;
  const int ty = tile % a.nty, tx = tile / a.nty;
  const int y = ty * WY - T + w;
  const bool row_valid = y >= 0 && y < a.Y;
  const int yc = min(max(y, 0), a.Y - 1);  // address-safe row
  const int xa = tx * a.LX;
  const int xb = min(xa + a.LX, a.X);
  const int xs = max(xa - T, 0);
  const int xe = xb + T - 2;  // level T reaches xb - 1
  const ZCoef<CPL> zc = make_zcoef<CPL>(a, zb);

  CarryT<T, CPL> s;
#pragma unroll
  for (int l = 0; l < T; ++l) {
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
#pragma unroll
      for (int k = 0; k < 3; ++k) s.em[l][k][c] = s.hm[l][k][c] = s.cbm[l][k][c] = 0.0f;
#pragma unroll
      for (int k = 0; k < 4; ++k) s.psm[l][k][c] = 0.0f;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) s.sigm[l][k] = 0.0f;
  }
  {
    const int o = (xs * a.Y + yc) * Z;
#pragma unroll
    for (int k = 0; k < 3; ++k) load_col<CPL, vec>(s.e0[k], a.Ein + (k * a.N + o), zb, Z, row_valid);
  }
  Plane2<CPL> cur, nxt;
  fetch2<CPL, vec>(cur, a, xs, yc, w == NW - 1, zb, row_valid, w >= 1, zc);
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    L.sE[0][0][w][0][c * 64 + lane] = s.e0[0][c];
    L.sE[0][0][w][1][c * 64 + lane] = s.e0[2][c];
  }
  lds_barrier();
  const Out2 O = make_out2(a);
  const bool lane_ok = zb < Z;
  int par = 0;
  for (int x = xs; x <= xe; ++x) {
    planeT<T, CPL, WY, vec>(a, O, L, s, cur, nxt, zc, x, xs, xa, xb, xe, w, y, yc, lane, zb,
                            row_valid, lane_ok, par);
    cur = nxt;
    par ^= 1;
  }
}

template <int T, int CPL, int WY, bool vec>
__global__ __launch_bounds__(64 * (WY + 2 * T - 1)) void k_stream_deep(Stream2Args a) {
  __shared__ LdsT<T, CPL, WY> L;
  // Tiles adjacent in y (which share ghost rows) on the same XCD, as in the
  // two-step kernel.
  int tile = blockIdx.x;
  if ((a.ntiles & 7) == 0) tile = (tile & 7) * (a.ntiles >> 3) + (tile >> 3);
  deep_tile<T, CPL, WY, vec>(a, tile, L);
}

// Batched form: persistent workgroups take (simulation, tile) items from a
// device counter, as k_stream2_batch does for step pairs.  Launch t advances
// every simulation from step T t to T t + T and reads the buffer set t & 1
// (args holds two views per simulation); a simulation past its last step is
// skipped, and a launch that runs past a simulation's last step writes only
// state that is never read (snapshots stop at its last step, sources are not
// read there).  Forward only: no captures.
template <int T, int CPL, int WY, bool vec>
__global__ __launch_bounds__(64 * (WY + 2 * T - 1)) void k_stream_deep_batch(
    const Stream2Args* __restrict__ args, const int* __restrict__ tile_start, int nsims,
    int total, int* __restrict__ counter, int step) {
  __shared__ LdsT<T, CPL, WY> L;
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
    Stream2Args a = args[2 * lo + ((step / T) & 1)];
    if (step >= a.nsteps) continue;
    a.n = step;
    a.flags = 0;
    for (int l = 1; l <= T; ++l)
      if (out_slot(a.rare, step + l - 1) >= 0) a.flags |= snap_flag(l);
    deep_tile<T, CPL, WY, vec>(a, t - tile_start[lo], L);
  }
}

// Depth T at one cell per lane (zz <= 64), within the store resources'
// 2^31-byte reach.
bool stream_deep_supported(const Geom& g, int T) {
  return (T == 3 || T == 4) && g.Z <= 64 && fits_rsrc(g);
}

// Whether the default kernel choice takes three steps per launch: at
// zz <= 64 when the 11-row tiles fill every CU even at 8-plane chunks.
// Measured against two steps per launch (1198 steps, Gcell-updates/s):
// 512x512x64 124.2 vs 119.8, 300x300x64 118.1 vs 112.1, but 128x128x64 64.7
// vs 75.4 and 64x64x64 16.8 vs 19.8 (too few tiles for the longer ghost
// start).
bool stream_deep_preferred(const Geom& g) {
  return stream_deep_supported(g, 3) &&
         (long long)((g.Y + 10) / 11) * ((g.X + 7) / 8) >= device_cus();
}

// Owned rows per tile: 16 waves at depth 3, 13 at depth 4 (whose registers
// spill at 16 waves: 4 steps per launch ran at 79 against 124 Gcell-updates/s
// for 3 at 512x512x64).
constexpr int kDeepWY3 = 11;  // stream_deep_preferred's tile count assumes 11
constexpr int kDeepWY4 = 6;

int deep_batch_rows() { return kDeepWY3; }

int deep_batch_resident() {
  static const int n = [] {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &b, reinterpret_cast<const void*>(&k_stream_deep_batch<3, 1, kDeepWY3, true>),
            64 * (kDeepWY3 + 5), 0) != hipSuccess || b < 1)
      b = 1;
    return b;
  }();
  return n;
}

void launch_deep_batch(const Stream2Args* args, const int* tile_start, int nsims, int total,
                       int* counter, int step, int grid, hipStream_t s) {
  hipLaunchKernelGGL((k_stream_deep_batch<3, 1, kDeepWY3, true>), dim3(grid),
                     dim3(64 * (kDeepWY3 + 5)), 0, s, args, tile_start, nsims, total, counter,
                     step);
}

template <int T, int CPL, int WY>
static void launch_deep(const Stream2Args& a, hipStream_t s) {
  hipLaunchKernelGGL((k_stream_deep<T, CPL, WY, true>), dim3(a.ntiles),
                     dim3(64 * (WY + 2 * T - 1)), 0, s, a);
}

// Runs steps [r.begin, r.end) T at a time; fewer than T left run on the
// two-step kernel (and its one-step tail), on the same buffer sets.
hipError_t run_stream_deep(const Geom& g, const pjz_fdtd_problem& p, const float* mask,
                           const Src& src, const OutWin& win, char* ws, const Layout& L,
                           const StepRange& r, int T, hipStream_t s) {
  if (!stream_deep_supported(g, T)) return hipErrorInvalidValue;
  const int WY = T == 3 ? kDeepWY3 : kDeepWY4;
  const int nty = (g.Y + WY - 1) / WY;
  const int LX = p.tile_x > 0 ? p.tile_x : grid_chunk_x(g.X, nty, device_cus(), T);
  Stream2Args a = make_args2(g, p, mask, src, win, ws, L, WY, LX, 0);
  Stream2Rare* rare = (Stream2Rare*)(ws + L.rare);
  hipLaunchKernelGGL(k_put_rare, dim3(1), dim3(1), 0, s, rare,
                     rare_of(p, win, r.capture, r.begin));
  a.rare = rare;
  float* psi = (float*)(ws + L.psi);
  const int P = g.X * g.Y * g.NP;
  float* E[2] = {(float*)(ws + L.E0), (float*)(ws + L.E1)};
  auto slot = [&](int step) {
    if (!win.out || step < p.out_start || step >= p.out_stop ||
        (step - p.out_start) % p.out_step != 0)
      return -1;
    return (step - p.out_start) / p.out_step;
  };
  int n = r.begin;
  for (; n + T <= r.end; n += T) {
    const int c = *r.cur;
    a.Ein = E[c];
    a.Eout = E[1 - c];
    a.psiIn = psi + (c ? PSI_SET : 0) * (long long)P;
    a.psiOut = psi + (c ? 0 : PSI_SET) * (long long)P;
    a.n = n;
    a.flags = r.capture ? CAPTURE : 0;
    for (int l = 1; l <= T; ++l)
      if (slot(n + l - 1) >= 0) a.flags |= snap_flag(l);
    if (T == 3) launch_deep<3, 1, kDeepWY3>(a, s);
    else launch_deep<4, 1, kDeepWY4>(a, s);
    *r.cur = 1 - c;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || n >= r.end) return e;
  StepRange tail{n, r.end, r.cur, r.capture ? r.capture + (long long)(n - r.begin) * 3 * g.N : nullptr};
  return run_stream2(g, p, mask, src, win, ws, L, tail, s);
}

}  // namespace pjz
