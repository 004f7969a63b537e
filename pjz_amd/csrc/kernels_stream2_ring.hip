// Two-step streaming stepper, two cells per lane, with its inputs staged
// through an LDS ring two planes deep.
//
// The schedule is k_stream2's (level 1 at plane x, level 2 at x-1; rows
// y0-2 .. y0+WY; three phases split by two LDS barriers).  What differs is
// how a plane's inputs arrive: k_stream2 prefetches the next plane into
// registers, so each wave has one plane of loads in flight, and at 168
// registers a lane there is no room for a second.  Here every input of a
// row (E at x+1, H, cb, the psi entries, and the top row's E of the row
// above) is copied by global_load_lds straight into LDS, which needs no
// registers: plane x+2 is requested while plane x is computed, so two planes
// are in flight.  On the forward kernel's access pattern alone (the same
// loads, stores and barriers with the arithmetic removed) that took a pass
// from 419 to 364 us at 512 x 512 x 128.
//
// LDS (8-row tiles, 11 waves): the ring, 2 x 11 x 9 columns of 128 floats
// (99 KB), psi 2 x 11 x 2 x 64 floats, the top row's pair of columns; the
// row exchange of k_stream2 with one slot per buffer instead of two (the
// writes that k_stream2 double-buffers move one phase later: the row below
// reads them after the next barrier, and nothing reads them in between).
// 156 KB in all, one workgroup per CU.
//
// The compiler does not see these copies (they are issued by inline asm), so
// the kernel waits for them itself: before reading plane x it waits until at
// most K vector memory operations are outstanding, K being the fewest that
// every wave issues after plane x's copies (the copies of plane x+1 and the
// unconditional stores of planes x-2 and x-1).  Loads the compiler does see
// (source field, snapshot schedule) wait for everything, which is correct
// and happens on a few planes only.
#include "stream2_parts.h"

namespace pjz {

namespace {

constexpr int kRingCols = 9;  // E x y z at x+1, H x y z at x, cb x y z at x
constexpr int kRingPsi = 32;  // psi entries per row the ring holds (lo + hi)

template <int WY>
struct RingLds {
  static constexpr int NW = WY + 3;
  // Row exchange (cell c of lane l at [c * 64 + l]).
  float sE1[NW][2][128];  // input Ex, Ez of plane x+1 (read by the row below next plane)
  float sH1[NW][2][128];  // level-1 Hx, Hz of plane x
  float sE2[NW][2][128];  // level-1 Ex, Ez of plane x (read by the row below next plane)
  float sH2[NW][2][128];  // level-2 Hx, Hz of plane x-1
  // Inputs, slot = plane parity; columns in z order.
  float ring[2][NW][kRingCols][128];
  float psr[2][NW][2][64];  // psi components 0|1 and 2|3, 32 entries each
  float upr[2][2][128];     // top row: input Ex, Ez of the row above
};

// Copies issued per plane by every wave (the top row issues one more).
constexpr int kRingCopies = 7;
// Stores every plane issues whatever its rows (E, H, and psi as pairs).
constexpr int kRingStores = 10;

// Requests plane p's inputs of row yc into ring slot `slot`.  p is within the
// domain; the E columns come from plane p + 1 (clamped: the consumer zeroes
// them past the domain).  Columns go two per 16-byte copy: lanes 0-31 the
// even column, lanes 32-63 the odd one.
template <int WY>
__device__ inline void ring_copy(RingLds<WY>& L, const Stream2Args& a, int slot, int w, int p,
                                 int yc, bool top, int lane) {
  const int N = a.N, Z = a.Z;
  const int oh = (p * a.Y + yc) * Z;
  const int oe = (min(p + 1, a.X - 1) * a.Y + yc) * Z;
  const int half = lane >> 5;
  // 4 cells a lane; heights below 128 (multiples of 4) repeat the last cells
  const int dz = min((lane & 31) * 4, Z - 4);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int k = 2 * j + half;
    const float* g = k < 3   ? a.Ein + (k * N + oe)
                     : k < 6 ? a.Ein + (a.hoff + (k - 3) * N + oh)
                             : a.cb + ((k - 6) * N + oh);
    // cb is read once per plane and row: nontemporal (0.3% faster)
    if (j == 3) dma16nt(g + dz, lds_addr(&L.ring[slot][w][2 * j][0]));
    else dma16(g + dz, lds_addr(&L.ring[slot][w][2 * j][0]));
  }
  if (lane < 32) dma16nt(a.cb + (2 * N + oh) + dz, lds_addr(&L.ring[slot][w][8][0]));
  // psi: lanes 0-31 component 2q, lanes 32-63 component 2q + 1
  const int e = lane & 31;
  const int pb = (p * a.Y + yc) * a.NP;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const float* g =
        a.NP > 0 ? a.psiIn + ((2 * q + half) * a.P + pb + min(e, a.NP - 1)) : a.Ein;
    dma4(g, lds_addr(&L.psr[slot][w][q][0]));
  }
  if (top) {
    const int ou = (p * a.Y + min(yc + 1, a.Y - 1)) * Z;
    dma16(a.Ein + ((half ? 2 : 0) * N + ou) + dz, lds_addr(&L.upr[slot][0][0]));
  }
}

// Plane x's inputs from ring slot `slot` into registers, zeroed where
// k_stream2's fetch would not load them.
template <int WY>
__device__ inline void ring_read(Plane2<2>& P, const RingLds<WY>& L, const Stream2Args& a,
                                 int slot, int w, int x, int yc, bool top, int zb, bool active,
                                 bool e_rows, const ZCoef<2>& zc) {
  const bool ae = active && x + 1 < a.X;
  const bool ac = active && e_rows;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const f32x2 te = *reinterpret_cast<const f32x2*>(&L.ring[slot][w][k][zb]);
    const f32x2 th = *reinterpret_cast<const f32x2*>(&L.ring[slot][w][3 + k][zb]);
    const f32x2 tc = *reinterpret_cast<const f32x2*>(&L.ring[slot][w][6 + k][zb]);
    P.e[k][0] = ae ? te.x : 0.0f;
    P.e[k][1] = ae ? te.y : 0.0f;
    P.h[k][0] = active ? th.x : 0.0f;
    P.h[k][1] = active ? th.y : 0.0f;
    P.cb[k][0] = ac ? tc.x : 0.0f;
    P.cb[k][1] = ac ? tc.y : 0.0f;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const bool ok = zc.pml[c] && active && (k < 2 || e_rows);
      const float v = L.psr[slot][w][k >> 1][(k & 1) * 32 + (zc.pml[c] ? zc.pidx[c] : 0)];
      P.ps[k][c] = ok ? v : 0.0f;
    }
  if (top) {
    const bool v = active && yc + 1 < a.Y;
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      const f32x2 t = *reinterpret_cast<const f32x2*>(&L.upr[slot][m][zb]);
      P.up[m][0] = v ? t.x : 0.0f;
      P.up[m][1] = v ? t.y : 0.0f;
    }
  } else {
    P.up[0][0] = P.up[0][1] = P.up[1][0] = P.up[1][1] = 0.0f;
  }
}

// One plane of the sweep (k_stream2's phases, single-slot row exchange).
template <int WY>
__device__ __forceinline__ void ring_plane(const Stream2Args& a, const Out2& O, RingLds<WY>& L,
                                           Carry2<2>& s, const Plane2<2>& cur,
                                           const ZCoef<2>& zc, int x, int xa, int w, int y,
                                           int lane, int zb, bool row_valid, bool lane_ok) {
  constexpr int CPL = 2, NW = WY + 3;
  const bool e_rows = w >= 1;
  const bool h2_rows = w >= 1 && w <= WY + 1;
  const bool own = w >= 2 && w <= WY + 1;
  const bool l1 = x < a.X;
  const bool l2 = x >= xa && x >= 1;
  const bool act = row_valid && l1;
  const int Z = a.Z;

  // ---- A: level-1 H at plane x (input E of the row above from sE1, written
  // there in phase B of plane x-1)
  float eup[2][CPL];
  if (w < NW - 1) {
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      eup[0][c] = L.sE1[w + 1][0][c * 64 + lane];
      eup[1][c] = L.sE1[w + 1][1][c * 64 + lane];
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
    h_row<CPL>(hn1, psH1, s.e0, cur.e, eup, cur.h, psH, zc, zb, Z, a.dt, act);
  }
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    L.sH1[w][0][c * 64 + lane] = hn1[0][c];
    L.sH1[w][1][c * 64 + lane] = hn1[2][c];
  }
  lds_barrier();

  // ---- B: input E of plane x+1 for the row below; level-1 E at plane x,
  // then level-2 H at plane x-1
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    L.sE1[w][0][c * 64 + lane] = cur.e[0][c];
    L.sE1[w][1][c * 64 + lane] = cur.e[2][c];
  }
  float en1[3][CPL], psE1[2][CPL];
  if (e_rows) {
    float hdn[2][CPL], hp[2][CPL], psE[2][CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      hdn[0][c] = L.sH1[w - 1][0][c * 64 + lane];
      hdn[1][c] = L.sH1[w - 1][1][c * 64 + lane];
      hp[0][c] = s.h1m[1][c];
      hp[1][c] = s.h1m[2][c];
      psE[0][c] = cur.ps[2][c];
      psE[1][c] = cur.ps[3][c];
    }
    e_row<CPL>(en1, psE1, s.e0, hn1, hdn, hp, cur.cb, cur.sig, psE, zc, a, a.n, x, y, zb, lane,
               act);
    if ((a.flags & (SNAP1 | CAPTURE)) && own && act && x >= xa && x < min(xa + a.LX, a.X))
      store_rare<CPL, true>(a, a.flags & SNAP1, a.flags & CAPTURE, a.n, x, y, zb, en1, s.e0);
  } else {
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      en1[0][c] = en1[1][c] = en1[2][c] = 0.0f;
      psE1[0][c] = psE1[1][c] = 0.0f;
    }
  }
  float hn2[3][CPL], psH2[2][CPL];
  if (h2_rows && l2) {
    float e2up[2][CPL], psHm[2][CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      e2up[0][c] = L.sE2[w + 1][0][c * 64 + lane];
      e2up[1][c] = L.sE2[w + 1][1][c * 64 + lane];
      psHm[0][c] = s.psm[0][c];
      psHm[1][c] = s.psm[1][c];
    }
    h_row<CPL>(hn2, psH2, s.e1m, en1, e2up, s.h1m, psHm, zc, zb, Z, a.dt, row_valid);
  } else {
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      hn2[0][c] = hn2[1][c] = hn2[2][c] = 0.0f;
      psH2[0][c] = psH2[1][c] = 0.0f;
    }
  }
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    L.sH2[w][0][c * 64 + lane] = hn2[0][c];
    L.sH2[w][1][c * 64 + lane] = hn2[2][c];
  }
  lds_barrier();

  // ---- C: level-1 E of plane x for the row below (read next plane);
  // level-2 E at plane x-1 (the result on owned rows)
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    L.sE2[w][0][c * 64 + lane] = en1[0][c];
    L.sE2[w][1][c * 64 + lane] = en1[2][c];
  }
  const bool st = own && l2 && row_valid && x - 1 >= xa;
  float en2[3][CPL], psE2[2][CPL];
  if (own && l2 && row_valid) {
    float hdn[2][CPL], psEm[2][CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      hdn[0][c] = L.sH2[w - 1][0][c * 64 + lane];
      hdn[1][c] = L.sH2[w - 1][1][c * 64 + lane];
      psEm[0][c] = s.psm[2][c];
      psEm[1][c] = s.psm[3][c];
    }
    e_row<CPL>(en2, psE2, s.e1m, hn2, hdn, s.h2p, s.cbm, s.sigm, psEm, zc, a, a.n + 1, x - 1, y,
               zb, lane, true);
    if (st && (a.flags & (SNAP2 | CAPTURE)))
      store_rare<CPL, true>(a, a.flags & SNAP2, a.flags & CAPTURE, a.n + 1, x - 1, y, zb, en2,
                            s.e1m);
  } else {
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      en2[0][c] = en2[1][c] = en2[2][c] = 0.0f;
      psE2[0][c] = psE2[1][c] = 0.0f;
    }
  }
  {
    // Every plane issues the same kRingStores stores (kOff drops a row's).
    const unsigned ob = st && lane_ok ? (unsigned)((((x - 1) * a.Y + y) * Z + zb) * 4) : kOff;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      bstore<CPL>(en2[k], O.e, ob, k * O.N4);
      bstore<CPL>(hn2[k], O.h, ob, k * O.N4);
    }
    const int pb = ((x - 1) * a.Y + y) * a.NP;
    if (a.pair) {
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

// Absorber values of row yc at plane x (wave-uniform scalar loads).
__device__ inline void ring_sig(float (&sig)[3], const Stream2Args& a, int x, int yc, bool ok) {
  const int xm = min(x, a.X - 1);
#pragma unroll
  for (int k = 0; k < 3; ++k) sig[k] = ok ? uniform_load(a.mask, (k * a.X + xm) * a.Y + yc) : 0.0f;
}

// One tile (WY rows x LX planes) of one simulation for steps a.n, a.n+1.
template <int WY>
__device__ __forceinline__ void ring_tile(const Stream2Args& a, int tile, RingLds<WY>& L) {
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
  const bool e_rows = w >= 1;
  const ZCoef<2> zc = make_zcoef<2>(a, zb);

  Carry2<2> s;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
#pragma unroll
    for (int k = 0; k < 3; ++k) s.e1m[k][c] = s.h1m[k][c] = s.cbm[k][c] = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) s.psm[k][c] = 0.0f;
    s.h2p[0][c] = s.h2p[1][c] = 0.0f;
  }
  s.sigm[0] = s.sigm[1] = s.sigm[2] = 0.0f;
  const int xs = max(xa - 2, 0);
  const int xlast = min(xb, a.X - 1);  // last plane with data to copy
  ring_copy<WY>(L, a, 0, w, min(xs, xlast), yc, top, lane);
  ring_copy<WY>(L, a, 1, w, min(xs + 1, xlast), yc, top, lane);
  {
    const int o = (xs * a.Y + yc) * Z;
#pragma unroll
    for (int k = 0; k < 3; ++k) load_col<2, true>(s.e0[k], a.Ein + (k * a.N + o), zb, Z, row_valid);
  }
  // Input Ex, Ez of the first plane for the row below's level-1 H.
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    L.sE1[w][0][c * 64 + lane] = s.e0[0][c];
    L.sE1[w][1][c * 64 + lane] = s.e0[2][c];
  }
  lds_barrier();
  const Out2 O = make_out2(a);
  const bool lane_ok = zb < Z;
  Plane2<2> cur;
  float nsig[3];
  ring_sig(cur.sig, a, xs, yc, row_valid && xs < a.X && e_rows);
  for (int x = xs; x <= xb; ++x) {
    const int it = x - xs;
    const int slot = it & 1;
    // Wait for plane x's copies: after them every wave has issued at least
    // the copies of plane x+1 and the stores of planes x-2 and x-1.
    if (it >= 2)
      asm volatile("s_waitcnt vmcnt(%0)" ::"n"(2 * kRingStores + kRingCopies) : "memory");
    else if (it == 1)
      asm volatile("s_waitcnt vmcnt(%0)" ::"n"(kRingStores + kRingCopies) : "memory");
    else
      asm volatile("s_waitcnt vmcnt(%0)" ::"n"(kRingCopies) : "memory");
    const bool active = row_valid && x < a.X;
    ring_read<WY>(cur, L, a, slot, w, x, yc, top, zb, active, e_rows, zc);
    // This wave's reads of the slot are done before plane x+2 overwrites it.
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    ring_copy<WY>(L, a, slot, w, min(x + 2, xlast), yc, top, lane);
    ring_sig(nsig, a, x + 1, yc, x + 1 <= xb && row_valid && x + 1 < a.X && e_rows);
    ring_plane<WY>(a, O, L, s, cur, zc, x, xa, w, y, lane, zb, row_valid, lane_ok);
#pragma unroll
    for (int k = 0; k < 3; ++k) cur.sig[k] = nsig[k];
  }
  // The last planes' copies (repeats of the last plane, never read) land
  // before the ring is reused or the workgroup ends.
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

template <int WY>
__global__ __launch_bounds__(64 * (WY + 3)) void k_stream2_ring(Stream2Args a) {
  __shared__ RingLds<WY> L;
  // Tiles adjacent in y (which share ghost rows) run on the same XCD.
  int tile = blockIdx.x;
  if ((a.ntiles & 7) == 0) tile = (tile & 7) * (a.ntiles >> 3) + (tile >> 3);
  ring_tile<WY>(a, tile, L);
}

// Batched form: persistent workgroups take (simulation, tile) items from the
// per-XCD queues, as k_stream2_batch does (args: two views per simulation,
// selected by the parity of the step pair).
template <int WY>
__global__ __launch_bounds__(64 * (WY + 3)) void k_stream2_ring_batch(
    const Stream2Args* __restrict__ args, const int* __restrict__ tile_start, int nsims,
    int total, int* __restrict__ counter, const int* __restrict__ step_base, int step_off) {
  __shared__ RingLds<WY> L;
  __shared__ int s_item, s_next;
  if (threadIdx.x == 0) s_next = 0;
  const int step =
      (step_base ? *(const __attribute__((address_space(4))) int*)step_base : 0) + step_off;
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
              (out_slot(a.rare, step + 1) >= 0 ? SNAP2 : 0) |
              (((rare_cptr)a.rare)->cap ? CAPTURE : 0);
    ring_tile<WY>(a, t - tile_start[lo], L);
  }
}

}  // namespace

// Heights that are multiples of 4 in (64, 128] with at most 32 PML cells,
// within the store resources' 2^31 bytes.
bool stream2_ring_supported(const Geom& g) {
  return fits_rsrc(g) && g.Z > 64 && g.Z <= 128 && g.Z % 4 == 0 && g.NP <= kRingPsi;
}

static constexpr int kRingWY = 8;

hipError_t launch_stream2_ring(const Stream2Args& a, hipStream_t s) {
  hipLaunchKernelGGL(k_stream2_ring<kRingWY>, dim3(a.ntiles), dim3(64 * (kRingWY + 3)), 0, s, a);
  return hipGetLastError();
}

int stream2_ring_batch_resident() {
  static const int n = [] {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &b, reinterpret_cast<const void*>(&k_stream2_ring_batch<kRingWY>), 64 * (kRingWY + 3),
            0) != hipSuccess || b < 1)
      b = 1;
    return b;
  }();
  return n;
}

void launch_stream2_ring_batch(const Stream2Args* args, const int* tile_start, int nsims,
                               int total, int* counter, const int* step_base, int step_off,
                               int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_stream2_ring_batch<kRingWY>, dim3(grid), dim3(64 * (kRingWY + 3)), 0, s,
                     args, tile_start, nsims, total, counter, step_base, step_off);
}

}  // namespace pjz
