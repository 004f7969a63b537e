// Streaming reverse step: one fused launch per reversed time step, the
// adjoint counterpart of the one-step streaming stepper.  The two halves of
// a reverse step (kernels_adjoint.hip gives them cell by cell) have the
// forward kernel's dependency shape mirrored:
//   lH update at (x, y, z) reads lK = cb lE at x+1, y+1, z+1 (like the
//     forward H update reads E there) and lpsiE at z, z+1;
//   lE update at (x, y, z) reads the new lH at x-1, y-1, z-1 (like the
//     forward E update reads H there) and the new psi_H cotangent at z, z-1.
// So the same sweep works: one wavefront per y row (whole z column, CPL cells
// per lane, z neighbours by lane shuffles), a workgroup of WY rows plus one
// ghost row below (its lH feeds row y0's lE update), rows exchanging lK and
// the new lH through LDS, x swept over a chunk with the new lH of plane x-1
// carried in registers and plane xa-1 recomputed as a ghost plane.  All of
// lE, lH and both psi cotangents ping-pong between two sets (ghost cells read
// values their owners rewrite in the same launch); G is updated in place by
// the owner only.
//
// Per cell and step it reads lE, cb, lH, D, G (+ the psi cotangents in PML
// cells, the absorber per column) and writes lE, lH, G: 96 B of the 168 B the
// two cell-by-cell kernels move.
#include <algorithm>
#include <vector>

#include "stream_common.h"

namespace pjz {

struct AdjStreamArgs {
  const float* lEin;
  const float* lHin;
  float* lEout;
  float* lHout;
  const float* luin;   // psi_H cotangent pair (x at +0, y at +P)
  float* luout;
  const float* lpin;   // psi_E cotangent pair
  float* lpout;
  float* G;
  const float* D;      // E^n - E^(n+1) of the step being reversed
  const float* cb;
  const float* zt;
  const float* mask;
  int X, Y, Z, lo, hi, NP, N, P;
  float dt;
  int LX, nty, ntiles;
  int nsteps;  // steps of the simulation (batched runs)
};

// Row exchange buffers (cell c of lane l at [c * 64 + l]: conflict-free
// banks), double-buffered by plane parity: lK of plane x+1 is published from
// the prefetch while plane x is computed, so one barrier per plane (between
// the lH and the lE updates) orders everything.
template <int CPL, int WY>
struct AdjLds {
  float sK[2][WY + 1][2][64 * CPL];  // lK x, z of a plane, per row
  float sH[2][WY + 1][2][64 * CPL];  // new lH x, z of a plane, per row
};

// Everything one wave reads for plane x of its row, fetched one plane ahead.
template <int CPL>
struct AdjPlane {
  float le[3][CPL];   // lE at x + 1
  float cb[3][CPL];   // cb at x + 1
  float lh[3][CPL];   // lH at x
  float lu[2][CPL];   // psi_H cotangent at x (PML cells)
  float lp[2][CPL];   // psi_E cotangent at x (PML cells)
  float d[3][CPL];    // D at x (owner cells)
  float g[3][CPL];    // G at x (owner cells)
  float kue[2][CPL];  // lE x, z at (x, y+1) (top row of the workgroup only)
  float kuc[2][CPL];  // cb x, z at (x, y+1): lK = kuc kue, formed where used so
                      // the fetch does not wait for its own loads
  float sig[3];       // absorber at (x, y) (owner cells)
};

template <int CPL, int WY, bool vec>
__device__ inline void adj_fetch(AdjPlane<CPL>& P, const AdjStreamArgs& a, int x, int y, int w,
                                 int zb, bool active, bool owner_cells, const bool (&pml)[CPL],
                                 const int (&pidx)[CPL]) {
  const int N = a.N, Y = a.Y, Z = a.Z;
  const int o = (x * Y + y) * Z;
  const bool nx = active && x + 1 < a.X;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    load_col<CPL, vec>(P.le[k], a.lEin + (k * N + o + Y * Z), zb, Z, nx);
    load_col<CPL, vec>(P.cb[k], a.cb + (k * N + o + Y * Z), zb, Z, nx);
    load_col<CPL, vec>(P.lh[k], a.lHin + (k * N + o), zb, Z, active);
    load_col<CPL, vec>(P.d[k], a.D + (k * N + o), zb, Z, owner_cells);
    load_col<CPL, vec>(P.g[k], a.G + (k * N + o), zb, Z, owner_cells);
  }
  if (w == WY) {
    const bool v = active && y + 1 < Y;
    load_col<CPL, vec>(P.kue[0], a.lEin + (o + Z), zb, Z, v);
    load_col<CPL, vec>(P.kue[1], a.lEin + (2 * N + o + Z), zb, Z, v);
    load_col<CPL, vec>(P.kuc[0], a.cb + (o + Z), zb, Z, v);
    load_col<CPL, vec>(P.kuc[1], a.cb + (2 * N + o + Z), zb, Z, v);
  }
  const int pb = (x * Y + y) * a.NP;
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    const bool ph = pml[c] && active;
    const int pi = pb + pidx[c];
    P.lu[0][c] = ph ? a.luin[pi] : 0.0f;
    P.lu[1][c] = ph ? a.luin[a.P + pi] : 0.0f;
    P.lp[0][c] = ph ? a.lpin[pi] : 0.0f;
    P.lp[1][c] = ph ? a.lpin[a.P + pi] : 0.0f;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k)
    P.sig[k] = owner_cells ? a.mask[(k * a.X + x) * Y + y] : 0.0f;
}

// One tile (WY rows x LX planes) of one reversed step.
template <int CPL, int WY, bool vec>
__device__ __forceinline__ void adj_tile(const AdjStreamArgs& a, int tile, AdjLds<CPL, WY>& S) {
  const int X = a.X, Y = a.Y, Z = a.Z, N = a.N;
  const float dt = a.dt;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // 0 = ghost row
  const int zb = lane * CPL;
  const int ty = tile % a.nty, tx = tile / a.nty;
  const int y0 = ty * WY;
  const int y = y0 - 1 + w;
  const bool row_valid = y >= 0 && y < Y;
  const int yc = min(max(y, 0), Y - 1);
  const bool owner = w >= 1 && row_valid;
  const int xa = tx * a.LX;
  const int xb = min(xa + a.LX, X);

  auto& sK = S.sK;
  auto& sH = S.sH;

  // z coefficients of the lane's cells and of their z+1 / z-1 neighbours.
  float ikE[CPL], ikEp[CPL], bE[CPL], cE[CPL], ikH[CPL], ikHm[CPL], bH[CPL], cH[CPL];
  bool pml[CPL];
  int pidx[CPL];
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    const int z = zb + c;
    const int zz = min(z, Z - 1);
    ikE[c] = a.zt[IKE * Z + zz];
    ikEp[c] = z + 1 < Z ? a.zt[IKE * Z + z + 1] : 0.0f;
    bE[c] = a.zt[BE * Z + zz];
    cE[c] = a.zt[CE * Z + zz];
    ikH[c] = a.zt[IKH * Z + zz];
    ikHm[c] = z >= 1 && z - 1 < Z ? a.zt[IKH * Z + z - 1] : 0.0f;
    bH[c] = a.zt[BH * Z + zz];
    cH[c] = a.zt[CH * Z + zz];
    pml[c] = z < Z && (z < a.lo || z >= Z - a.hi);
    pidx[c] = z < a.lo ? z : z - (Z - a.hi) + a.lo;
  }

  // Carried: lE and cb at x (this row), new lH y, z at x-1.
  float le0[3][CPL], cb0[3][CPL], hyp[CPL], hzp[CPL];
  const int xs = max(xa - 1, 0);
  {
    const int o = (xs * Y + yc) * Z;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      load_col<CPL, vec>(le0[k], a.lEin + (k * N + o), zb, Z, row_valid);
      load_col<CPL, vec>(cb0[k], a.cb + (k * N + o), zb, Z, row_valid);
    }
#pragma unroll
    for (int c = 0; c < CPL; ++c) hyp[c] = hzp[c] = 0.0f;
  }
  AdjPlane<CPL> cur, nxt;
  adj_fetch<CPL, WY, vec>(cur, a, xs, yc, w, zb, row_valid, owner && xs >= xa, pml, pidx);
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    sK[0][w][0][c * 64 + lane] = cb0[0][c] * le0[0][c];
    sK[0][w][1][c * 64 + lane] = cb0[2][c] * le0[2][c];
  }
  lds_barrier();

  for (int x = xs; x < xb; ++x) {
    const bool ghost_plane = x < xa;
    const bool store = owner && !ghost_plane;
    const int o = (x * Y + y) * Z;
    adj_fetch<CPL, WY, vec>(nxt, a, x + 1, yc, w, zb, row_valid && x + 1 < xb,
                            owner && x + 1 < xb, pml, pidx);
    float K[3][CPL];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int c = 0; c < CPL; ++c) K[k][c] = cb0[k][c] * le0[k][c];
    const int par = (x - xs) & 1;
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      sK[par ^ 1][w][0][c * 64 + lane] = cur.cb[0][c] * cur.le[0][c];
      sK[par ^ 1][w][1][c * 64 + lane] = cur.cb[2][c] * cur.le[2][c];
    }
    float kyp[2][CPL];  // lK x, z of the row above
    if (w < WY) {
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        kyp[0][c] = sK[par][w + 1][0][c * 64 + lane];
        kyp[1][c] = sK[par][w + 1][1][c * 64 + lane];
      }
    } else {
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        kyp[0][c] = cur.kuc[0][c] * cur.kue[0][c];
        kyp[1][c] = cur.kuc[1][c] * cur.kue[1][c];
      }
    }

    // ---- lH update at plane x, row y
    float kxz[CPL], kyz[CPL], vx[CPL], vy[CPL], vxz[CPL], vyz[CPL], h[3][CPL], lun[2][CPL];
    z_next<CPL>(kxz, K[0], zb, Z);
    z_next<CPL>(kyz, K[1], zb, Z);
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      vx[c] = pml[c] ? cE[c] * (cur.lp[0][c] - K[0][c]) : 0.0f;
      vy[c] = pml[c] ? cE[c] * (cur.lp[1][c] + K[1][c]) : 0.0f;
    }
    z_next<CPL>(vxz, vx, zb, Z);
    z_next<CPL>(vyz, vy, zb, Z);
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      // lK y, z at x+1 from the prefetched plane.
      const float kxp_y = cur.cb[1][c] * cur.le[1][c];
      const float kxp_z = cur.cb[2][c] * cur.le[2][c];
      h[0][c] = cur.lh[0][c] - (ikEp[c] * kyz[c] - ikE[c] * K[1][c]) + (kyp[1][c] - K[2][c]) -
                (vyz[c] - vy[c]);
      h[1][c] = cur.lh[1][c] + (ikEp[c] * kxz[c] - ikE[c] * K[0][c]) - (kxp_z - K[2][c]) -
                (vxz[c] - vx[c]);
      h[2][c] = cur.lh[2][c] - (kyp[0][c] - K[0][c]) + (kxp_y - K[1][c]);
      // Rows outside the domain hold no cells.
      if (!row_valid) h[0][c] = h[1][c] = h[2][c] = 0.0f;
      lun[0][c] = pml[c] ? bH[c] * cur.lu[0][c] + dt * h[0][c] : 0.0f;
      lun[1][c] = pml[c] ? bH[c] * cur.lu[1][c] - dt * h[1][c] : 0.0f;
    }
    if (store) {
#pragma unroll
      for (int k = 0; k < 3; ++k) store_col<CPL, vec>(a.lHout + (k * N + o), h[k], zb, Z);
      const int pb = (x * Y + y) * a.NP;
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        if (pml[c]) {
          a.luout[pb + pidx[c]] = lun[0][c];
          a.luout[a.P + pb + pidx[c]] = lun[1][c];
        }
      }
    }
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      sH[par][w][0][c * 64 + lane] = h[0][c];
      sH[par][w][1][c * 64 + lane] = h[2][c];
    }
    lds_barrier();

    // ---- lE update at plane x, row y (owner rows, chunk planes)
    if (store) {
      float hxz[CPL], hyz[CPL], ux[CPL], uy[CPL], uxz[CPL], uyz[CPL];
      z_prev<CPL>(hxz, h[0], lane);
      z_prev<CPL>(hyz, h[1], lane);
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        ux[c] = pml[c] ? cH[c] * lun[0][c] : 0.0f;
        uy[c] = pml[c] ? cH[c] * lun[1][c] : 0.0f;
      }
      z_prev<CPL>(uxz, ux, lane);
      z_prev<CPL>(uyz, uy, lane);
      float len[3][CPL], gn[3][CPL];
      const int pb = (x * Y + y) * a.NP;
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        const float hx_ym = sH[par][w - 1][0][c * 64 + lane];
        const float hz_ym = sH[par][w - 1][1][c * 64 + lane];
        const float gx = dt * ((ikH[c] * h[1][c] - ikHm[c] * hyz[c]) - (h[2][c] - hz_ym)) -
                         (uy[c] - uyz[c]);
        const float gy = dt * (-(ikH[c] * h[0][c] - ikHm[c] * hxz[c]) + (h[2][c] - hzp[c])) -
                         (ux[c] - uxz[c]);
        const float gz = dt * ((h[0][c] - hx_ym) - (h[1][c] - hyp[c]));
        const float gs[3] = {gx, gy, gz};
        float lk[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const float le = le0[k][c];
          gn[k][c] = cur.g[k][c] + le * cur.d[k][c];
          lk[k] = cb0[k][c] * le;
          len[k][c] = (1.0f - cur.sig[k] * cb0[k][c]) * le + gs[k];
        }
        if (pml[c]) {
          a.lpout[pb + pidx[c]] = bE[c] * (cur.lp[0][c] - lk[0]);
          a.lpout[a.P + pb + pidx[c]] = bE[c] * (cur.lp[1][c] + lk[1]);
        }
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        store_col<CPL, vec>(a.lEout + (k * N + o), len[k], zb, Z);
        store_col<CPL, vec>(a.G + (k * N + o), gn[k], zb, Z);
      }
    }
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      hyp[c] = h[1][c];
      hzp[c] = h[2][c];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        le0[k][c] = cur.le[k][c];
        cb0[k][c] = cur.cb[k][c];
      }
    }
    cur = nxt;
  }
}

template <int CPL, int WY, bool vec>
__global__ __launch_bounds__(64 * (WY + 1)) void k_adj_stream(AdjStreamArgs a) {
  __shared__ AdjLds<CPL, WY> S;
  // Tiles adjacent in y (sharing ghost rows) go to the same XCD.
  int tile = blockIdx.x;
  if ((a.ntiles & 7) == 0) tile = (tile & 7) * (a.ntiles >> 3) + (tile >> 3);
  adj_tile<CPL, WY, vec>(a, tile, S);
}

// Batched reverse step: persistent workgroups take (simulation, tile) items
// from a device counter, as the batched two-step forward kernel does.  args
// holds two views per simulation (which adjoint set is read); a simulation
// that joined the backward at launch lstart[s] reads set
// (launch - lstart[s]) & 1.  D of the step sits at slot step - seg_begin of
// the view's D base.
template <int CPL, int WY, bool vec>
__global__ __launch_bounds__(64 * (WY + 1)) void k_adj_stream_batch(
    const AdjStreamArgs* __restrict__ args, const int* __restrict__ tile_start,
    const int* __restrict__ lstart, int nsims, int total, int* __restrict__ counter, int launch,
    int step, int seg_begin) {
  __shared__ AdjLds<CPL, WY> S;
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
    AdjStreamArgs a = args[2 * lo + ((launch - lstart[lo]) & 1)];
    a.D += (long long)(step - seg_begin) * 3 * a.N;
    adj_tile<CPL, WY, vec>(a, t - tile_start[lo], S);
  }
}

// Whether the streaming reverse step handles this problem.
bool adjoint_stream_supported(const Geom& g) {
  return (g.Z <= 64) || (g.Z <= 128 && g.Z % 2 == 0) || (g.Z <= 256 && g.Z % 4 == 0);
}

static int device_cus() {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  return cus;
}

static int rows_per_tile(const Geom& g) { return g.Z > 64 && g.Z <= 128 ? 8 : 4; }

// Arguments of one reversed step reading adjoint set acur; LX = 0 picks x
// chunks that give about one workgroup per CU.
static AdjStreamArgs adj_stream_args(const Geom& g, const pjz_fdtd_problem& p, const float* mask,
                                     char* ws, const Layout& L, const float* D, int acur, int LX) {
  AdjStreamArgs a;
  const size_t in = acur ? 1 : 0, out = 1 - in;
  a.lEin = (const float*)(ws + L.lE_[in]);
  a.lHin = (const float*)(ws + L.lH_[in]);
  a.luin = (const float*)(ws + L.lu_[in]);
  a.lpin = (const float*)(ws + L.lpE_[in]);
  a.lEout = (float*)(ws + L.lE_[out]);
  a.lHout = (float*)(ws + L.lH_[out]);
  a.luout = (float*)(ws + L.lu_[out]);
  a.lpout = (float*)(ws + L.lpE_[out]);
  a.G = (float*)(ws + L.accG);
  a.D = D;
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
  const int WY = rows_per_tile(g);
  a.nty = (g.Y + WY - 1) / WY;
  if (LX <= 0) {
    LX = grid_chunk_x(g.X, a.nty, device_cus(), 1);
  }
  a.LX = LX;
  a.ntiles = a.nty * ((g.X + a.LX - 1) / a.LX);
  a.nsteps = steps_of(p);
  return a;
}

hipError_t launch_adjoint_stream(const Geom& g, const pjz_fdtd_problem& p, const float* mask,
                                 char* ws, const Layout& L, const float* D, int acur,
                                 hipStream_t s) {
  if (!adjoint_stream_supported(g)) return hipErrorInvalidValue;
  AdjStreamArgs a = adj_stream_args(g, p, mask, ws, L, D, acur, 0);
  const int cpl = g.Z <= 64 ? 1 : (g.Z <= 128 ? 2 : 4);
  if (cpl == 1)
    hipLaunchKernelGGL((k_adj_stream<1, 4, true>), dim3(a.ntiles), dim3(64 * 5), 0, s, a);
  else if (cpl == 2)
    hipLaunchKernelGGL((k_adj_stream<2, 8, true>), dim3(a.ntiles), dim3(64 * 9), 0, s, a);
  else
    hipLaunchKernelGGL((k_adj_stream<4, 4, true>), dim3(a.ntiles), dim3(64 * 5), 0, s, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------- batches
// Meta layout: two argument views per simulation (cells-per-lane class
// order) | tile prefix sums (n_c + 1 per class) | queue counters (2 x
// kQueues per reversed step of a segment).
static constexpr int kAdjBatchLX = 128;

size_t adjoint_batch_meta_bytes(int n, int S) {
  return align256(sizeof(AdjStreamArgs) * 2 * (size_t)n) + align256(sizeof(int) * (size_t)(n + 2)) +
         align256(sizeof(int) * (size_t)n) + align256(sizeof(int) * 2 * kQueues * (size_t)std::max(S, 1));
}

hipError_t adjoint_batch_prepare(const BatchItem* items, int n, int S, const int* lstart,
                                 char* meta, AdjBatchPlan& plan, hipStream_t s) {
  std::vector<int> cls[2];
  for (int i = 0; i < n; ++i) {
    if (!adjoint_stream_supported(items[i].g) || items[i].g.Z > 128) return hipErrorInvalidValue;
    cls[items[i].g.Z <= 64 ? 0 : 1].push_back(i);
  }
  plan.n = n;
  plan.S = std::max(S, 1);
  plan.meta = meta;
  plan.cus = device_cus();
  plan.args_bytes = align256(sizeof(AdjStreamArgs) * 2 * (size_t)n);
  plan.ts_bytes = align256(sizeof(int) * (size_t)(n + 2));
  plan.ls_bytes = align256(sizeof(int) * (size_t)n);
  std::vector<char> host(plan.args_bytes + plan.ts_bytes + plan.ls_bytes, 0);
  AdjStreamArgs* hargs = (AdjStreamArgs*)host.data();
  int* hts = (int*)(host.data() + plan.args_bytes);
  int* hls = (int*)(host.data() + plan.args_bytes + plan.ts_bytes);
  int off = 0;
  for (int c = 0; c < 2; ++c) {
    plan.base[c] = off;
    plan.count[c] = (int)cls[c].size();
    // x chunks from the same greedy-queue model as the forward batch, on the
    // one workgroup per CU this kernel's launch uses.
    std::vector<int> xs, ntys;
    for (int i : cls[c]) {
      xs.push_back(items[i].g.X);
      ntys.push_back((items[i].g.Y + rows_per_tile(items[i].g) - 1) / rows_per_tile(items[i].g));
    }
    const int cLX = xs.empty() ? kAdjBatchLX
                               : batch_chunk_x(xs.data(), ntys.data(), (int)xs.size(), plan.cus,
                                               kAdjBatchLX);
    int t = 0;
    for (size_t k = 0; k < cls[c].size(); ++k) {
      const BatchItem& it = items[cls[c][k]];
      const int LX = std::min(cLX, it.g.X);
      for (int v = 0; v < 2; ++v)
        hargs[2 * (off + k) + v] =
            adj_stream_args(it.g, *it.p, it.mask, it.ws, *it.L, it.capture, v, LX);
      hts[off + c + k] = t;
      hls[off + k] = lstart[cls[c][k]];
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

hipError_t adjoint_batch_begin_segment(const AdjBatchPlan& plan, hipStream_t s) {
  int* counters = (int*)(plan.meta + plan.args_bytes + plan.ts_bytes + plan.ls_bytes);
  return launch_zero((float*)counters, 2 * kQueues * (size_t)plan.S, s);
}

hipError_t launch_adjoint_batch_step(const AdjBatchPlan& plan, int launch, int step,
                                     int seg_begin, hipStream_t s) {
  const int slot = step - seg_begin;
  if (slot < 0 || slot >= plan.S) return hipErrorInvalidValue;
  const AdjStreamArgs* dargs = (const AdjStreamArgs*)plan.meta;
  const int* dts = (const int*)(plan.meta + plan.args_bytes);
  const int* dls = (const int*)(plan.meta + plan.args_bytes + plan.ts_bytes);
  int* counters = (int*)(plan.meta + plan.args_bytes + plan.ts_bytes + plan.ls_bytes);
  for (int c = 0; c < 2; ++c) {
    const int nc = plan.count[c];
    if (nc == 0) continue;
    const int grid = std::min(plan.total[c], plan.cus);
    const AdjStreamArgs* a = dargs + 2 * plan.base[c];
    const int* ts = dts + plan.base[c] + c;
    const int* ls = dls + plan.base[c];
    int* ctr = counters + (2 * slot + c) * kQueues;
    if (c == 0)
      hipLaunchKernelGGL((k_adj_stream_batch<1, 4, true>), dim3(grid), dim3(64 * 5), 0, s, a, ts,
                         ls, nc, plan.total[c], ctr, launch, step, seg_begin);
    else
      hipLaunchKernelGGL((k_adj_stream_batch<2, 8, true>), dim3(grid), dim3(64 * 9), 0, s, a, ts,
                         ls, nc, plan.total[c], ctr, launch, step, seg_begin);
  }
  return hipGetLastError();
}

}  // namespace pjz
