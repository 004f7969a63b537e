// Shared definitions of the gfx950 Yee FDTD library: problem geometry,
// workspace layout, the per-z CPML coefficient table and launch plumbing.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "../../include/pjz_fdtd.h"

namespace pjz {

// Rows of the per-z coefficient table (zt), each `zz` floats long.
//   IKE/BE/CE: 1/kappa, b, c of pml column 0 (E update, z = k)
//   IKH/BH/CH: 1/kappa, b, c of pml column 1 (H update, z = k + 1/2)
enum ZRow { IKE = 0, BE = 1, CE = 2, IKH = 3, BH = 4, CH = 5, ZROWS = 8 };

// psi components, each (xx, yy, np) with np = pml_lo + pml_hi:
//   PSI_HX: d/dz Ey in the Hx update   PSI_HY: d/dz Ex in the Hy update
//   PSI_EX: d/dz Hy in the Ex update   PSI_EY: d/dz Hx in the Ey update
// in two sets of PSI_SET components.  The streaming kernels recompute cells
// owned by other workgroups (ghost rows and planes), so they read psi from one
// set and write the other, like E and H; the reference stepper uses set 0 in
// place.
enum PsiComp { PSI_HX = 0, PSI_HY = 1, PSI_EX = 2, PSI_EY = 3, PSI_SET = 4, PSI_N = 8 };

struct Geom {
  int X, Y, Z;      // domain
  int lo, hi, NP;   // pml widths, NP = lo + hi
  long long N;      // X*Y*Z
};

// Source description handed to the E update.
struct Src {
  int axis, pos, quad;
  const float* field;
  const float* wave;  // (tt, 2)
};

// Output window of the snapshots.
struct OutWin {
  int x0, y0, z0, ex, ey, ez;
  float* out;  // (nn, 3, ex, ey, ez)
};

// Byte offsets of the workspace segments (each 256-byte aligned).
struct Layout {
  size_t E0, H0, E1, H1, cb, psi, zt;
  size_t rare;        // 256 B of per-call launch data for the two-step kernel
  // Gradient support (only when ckpt_every > 0):
  size_t ckpt;        // nck checkpoints of ckpt_floats each
  size_t seg;         // (S + 1) x 3N: E^n - E^(n+1) of every step of a segment (+1 spare)
  // Adjoint state in two ping-pong sets (the streaming reverse step reads one
  // and writes the other; the cell-by-cell one works in place):
  size_t lE_[2], lH_[2];   // adjoint E and H (3N each)
  size_t lu_[2], lpE_[2];  // psi_H and psi_E cotangents (2P each)
  size_t accG;        // gradient accumulator sum_n lE^(n+1) (E^n - E^(n+1)) (3N)
  size_t total;
  int S, nck;
  // Two-level checkpoints: the forward keeps nck states every S1 steps
  // (slots 0 .. nck-1); the backward keeps up to nfine states every S steps
  // of one S1 segment in slots nck .. nck+nfine-1.  One level: S1 = S,
  // nfine = 0.
  int S1, nfine;
  size_t ckpt_floats;  // E (3N) + H (3N) + psi_H pair (2P) + psi_E pair (2P)
};

inline size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// Number of FDTD steps a call runs: the last snapshot step + 1.
inline int steps_of(const pjz_fdtd_problem& p) {
  if (p.out_stop <= p.out_start) return 0;
  int nn = (p.out_stop - p.out_start + p.out_step - 1) / p.out_step;
  return p.out_start + (nn - 1) * p.out_step + 1;
}

inline Layout make_layout(const Geom& g, const pjz_fdtd_problem& p) {
  Layout L;
  const size_t P = (size_t)g.X * g.Y * (size_t)(g.NP > 0 ? g.NP : 1);
  const size_t f3 = align256(sizeof(float) * 3 * (size_t)g.N);
  const size_t p2 = align256(sizeof(float) * 2 * P);
  size_t off = 0;
  L.E0 = off; off += f3;
  L.H0 = off; off += f3;
  L.E1 = off; off += f3;
  L.H1 = off; off += f3;
  L.cb = off; off += f3;
  L.psi = off;
  off += align256(sizeof(float) * PSI_N * P);
  L.zt = off; off += align256(sizeof(float) * ZROWS * (size_t)g.Z);
  L.rare = off; off += 256;
  L.S = 0;
  L.nck = 0;
  L.S1 = 0;
  L.nfine = 0;
  L.ckpt_floats = 6 * (size_t)g.N + 4 * P;
  L.ckpt = L.seg = L.accG = off;
  for (int i = 0; i < 2; ++i) L.lE_[i] = L.lH_[i] = L.lu_[i] = L.lpE_[i] = off;
  const int nsteps = steps_of(p);
  if (p.ckpt_every > 0 && nsteps > 0) {
    L.S1 = p.ckpt_every < nsteps ? p.ckpt_every : nsteps;
    L.nck = (nsteps + L.S1 - 1) / L.S1;
    L.S = L.S1;
    if (p.ckpt_inner > 0 && p.ckpt_inner < L.S1) {
      L.S = p.ckpt_inner;
      L.nfine = (L.S1 + L.S - 1) / L.S;
    }
    L.ckpt = off;
    off += align256(sizeof(float) * L.ckpt_floats) * (size_t)(L.nck + L.nfine);
    // D^b .. D^(b+S-1) of a segment, plus one slot for the extra step a
    // batched pair runs past an odd step count.
    L.seg = off; off += f3 * (size_t)(L.S + 1);
    for (int i = 0; i < 2; ++i) {
      L.lE_[i] = off; off += f3;
      L.lH_[i] = off; off += f3;
      L.lu_[i] = off; off += p2;
      L.lpE_[i] = off; off += p2;
    }
    L.accG = off; off += f3;
  }
  L.total = off;
  return L;
}

inline size_t ckpt_stride(const Layout& L) { return align256(sizeof(float) * L.ckpt_floats); }

// A run of steps [begin, end).  *cur names the ping-pong buffer set that holds
// the state before the run and is updated to the set that holds it after (the
// reference stepper works in place on set 0).  If capture is set, the change
// of E over step n, D^n = E^n - E^(n+1), of the whole domain is also written
// to capture + (n-begin)*3N (what the gradient needs of the forward fields).
struct StepRange {
  int begin, end;
  int* cur;
  float* capture;
};

// Launchers implemented in the kernel translation units.
hipError_t launch_prepare(const Geom& g, const pjz_fdtd_problem& p, const float* eps,
                          const float* mask, const float* kappa, const float* sigma,
                          const float* alpha, float* cb, float* zt, hipStream_t s);

// Clears n floats with a fill kernel (no hipMemsetAsync, so the call can sit
// inside a captured graph).
hipError_t launch_zero(float* ptr, size_t n, hipStream_t s);

// Copies n floats with a copy kernel.
hipError_t launch_copy(float* dst, const float* src, size_t n, hipStream_t s);

// Copies n16 16-byte words with nontemporal loads and stores (bandwidth probe).
hipError_t launch_copy_stream(void* dst, const void* src, size_t n16, hipStream_t s);

// x chunk length for a batch class: minimises rounds x planes per tile over
// candidate lengths <= max_lx (xs, nty: each simulation's X and row tiles;
// `ghost`: recomputed planes per tile).
int batch_chunk_x(const int* xs, const int* nty, int n, int slots, int max_lx, int ghost = 2);
int grid_chunk_x(int X, int nty, int slots, int ghost);

// Waveguide mode solver (kernels_mode.hip): ww frequencies of a (xx, yy)
// cross-section, 1 <= mm <= kModeMaxModes modes, one workgroup per frequency.
constexpr int kModeMaxModes = 10;
size_t mode_workspace_floats(int ww, int xx, int yy, int mm);
hipError_t launch_mode(int ww, int xx, int yy, int mm, const float* eps, const float* omega,
                       const float* shift, float* x, float* w, float* err, int* iters,
                       int max_iters, float tol, float* ws, hipStream_t s);

hipError_t run_reference(const Geom& g, const pjz_fdtd_problem& p, const float* mask,
                         const Src& src, const OutWin& win, char* ws, const Layout& L,
                         const StepRange& r, hipStream_t s);

hipError_t run_stream(const Geom& g, const pjz_fdtd_problem& p, const float* mask,
                      const Src& src, const OutWin& win, char* ws, const Layout& L,
                      const StepRange& r, hipStream_t s);

// Two steps per launch (temporal blocking of depth 2); an odd last step uses
// run_stream.  stream2_supported: whether the problem's zz allows it.
bool stream2_supported(const Geom& g);
hipError_t run_stream2(const Geom& g, const pjz_fdtd_problem& p, const float* mask,
                       const Src& src, const OutWin& win, char* ws, const Layout& L,
                       const StepRange& r, hipStream_t s);

// T = 3 or 4 steps per launch (deeper temporal blocking, zz <= 64); fewer
// than T remaining steps run on run_stream2.
bool stream_deep_supported(const Geom& g, int T);
// Whether the default (AUTO) kernel choice is three steps per launch.
bool stream_deep_preferred(const Geom& g);
hipError_t run_stream_deep(const Geom& g, const pjz_fdtd_problem& p, const float* mask,
                           const Src& src, const OutWin& win, char* ws, const Layout& L,
                           const StepRange& r, int T, hipStream_t s);

// One simulation of a batch: its problem, arrays and workspace.  `capture`
// (or null) receives E of the whole domain after each step, as in StepRange,
// counted from the step set by batch2_capture_from.
struct BatchItem {
  Geom g;
  const pjz_fdtd_problem* p;
  const float* mask;
  Src src;
  OutWin win;
  char* ws;
  const Layout* L;
  float* capture;
};

// Batched two-step stepping.  The items' workspaces must be prepared
// (coefficients, cleared state); `meta` is device scratch of
// batch2_meta_bytes() for argument views, tile tables and queue counters.
// Pair m advances every simulation from step 2m to 2m+2 and reads the buffer
// set m & 1 (a simulation skips pairs past its last step).
struct Batch2Plan {
  int n, npairs, cus;
  int chunk;               // step pairs per replayed graph chunk (0: direct launches)
  hipGraph_t graph;        // captured chunk, built on first use
  hipGraphExec_t exec;
  int count[2], base[2], total[2];  // per cells-per-lane class: sims, first sim, tiles
  int wy[2];                        // per class: rows per tile
  int depth0, nlaunch0;  // zz <= 64 class: steps per launch (2, or 3 in
                         // forward-only batches) and its three-step launches
  bool ring1;            // zz > 64 class: every simulation fits the LDS-ring kernel
  char* meta;
  size_t args_bytes, rare_bytes, ts_bytes;
};
bool batch2_supported(const Geom& g);
size_t batch2_meta_bytes(const BatchItem* items, int n);
hipError_t batch2_prepare(const BatchItem* items, int n, char* meta, Batch2Plan& plan,
                          hipStream_t s, bool forward_only = false);
hipError_t batch2_pairs(Batch2Plan& plan, int m0, int m1, hipStream_t s);
// Every step of a forward-only plan (batch2_prepare with forward_only): the
// zz <= 64 class may run three steps per launch; the final state is left in
// either buffer set (only snapshots are read).
hipError_t batch2_forward(Batch2Plan& plan, hipStream_t s);
// Frees the plan's graph (synchronising the stream if one was built).
hipError_t batch2_release(Batch2Plan& plan, hipStream_t s);
hipError_t batch2_capture_from(const Batch2Plan& plan, int begin, hipStream_t s);

// Reverse of step n: adjoint set `acur` (lE, lH, lu, lpE) holds the
// cotangents after step n on entry and before it on exit (in place); G
// accumulates lE' D^n (D^n = E^n - E^(n+1)).
hipError_t launch_adjoint_step(const Geom& g, const pjz_fdtd_problem& p, const float* mask,
                               char* ws, const Layout& L, const float* D, int acur,
                               hipStream_t s);

// Two reversed steps n, n-1 (D1 = D^n, D2 = D^(n-1)) in one launch: reads
// adjoint set `acur`, writes set 1 - acur.
bool adjoint_stream2_supported(const Geom& g);
hipError_t launch_adjoint_stream2(const Geom& g, const pjz_fdtd_problem& p, const float* mask,
                                  char* ws, const Layout& L, const float* D1, const float* D2,
                                  int acur, hipStream_t s);

// The same reverse step as one streaming launch: reads adjoint set `acur`,
// writes set 1 - acur.
bool adjoint_stream_supported(const Geom& g);
hipError_t launch_adjoint_stream(const Geom& g, const pjz_fdtd_problem& p, const float* mask,
                                 char* ws, const Layout& L, const float* D, int acur,
                                 hipStream_t s);

// Batched reverse steps: one persistent streaming launch per cells-per-lane
// class for every simulation of a batch, (simulation, tile) items taken from
// a device queue.  Each item's D^n is in its `capture` buffer at slot
// n - seg_begin; a simulation skips steps at or past its last.  The backward
// is a fixed sequence of launches (one or two reversed steps each); a
// simulation that joins at launch lstart[i] reads adjoint set
// (launch - lstart[i]) & 1.  S bounds the steps of a segment;
// *_begin_segment clears the queue counters.
struct AdjBatchPlan {
  int n, S, cus;
  int count[2], base[2], total[2];
  char* meta;
  size_t args_bytes, ts_bytes, ls_bytes;
  bool ring1;  // two-step reverse: the zz > 64 class fits the LDS-ring kernel
};
size_t adjoint_batch_meta_bytes(int n, int S);
hipError_t adjoint_batch_prepare(const BatchItem* items, int n, int S, const int* lstart,
                                 char* meta, AdjBatchPlan& plan, hipStream_t s);
hipError_t adjoint_batch_begin_segment(const AdjBatchPlan& plan, hipStream_t s);
hipError_t launch_adjoint_batch_step(const AdjBatchPlan& plan, int launch, int step,
                                     int seg_begin, hipStream_t s);

// Pairs of reversed steps (step, step-1) for a batch (the two-step reverse
// kernel in persistent-queue form); same conventions.
using Adj2BatchPlan = AdjBatchPlan;
size_t adjoint2_batch_meta_bytes(int n, int S);
hipError_t adjoint2_batch_prepare(const BatchItem* items, int n, int S, const int* lstart,
                                  char* meta, Adj2BatchPlan& plan, hipStream_t s);
hipError_t adjoint2_batch_begin_segment(const Adj2BatchPlan& plan, hipStream_t s);
hipError_t launch_adjoint2_batch_step(const Adj2BatchPlan& plan, int launch, int step,
                                      int seg_begin, hipStream_t s);

// pjz's density() on nb (X, Y) variables (kernels_density.hip): the cone's
// half-width r = ceil(radius - 1/2); forward writes density, loss and the
// blurred variable uf (the backward's input); backward needs a workspace of
// density_workspace_bytes.
int density_half_width(float radius);
size_t density_workspace_bytes(int nb, int X, int Y, float radius);
hipError_t density_forward(int nb, int X, int Y, float radius, float alpha, float c, float eta,
                           float lo, float hi, const float* u, float* dens, float* loss,
                           float* uf, hipStream_t s);
hipError_t density_backward(int nb, int X, int Y, float radius, float alpha, float c, float eta,
                            float lo, float hi, const float* uf, const float* gd,
                            const float* gl, float* gu, char* ws, hipStream_t s);

// pjz's render (kernels_render.hip): ll layers (ll, 2m xx, 2m yy), ll - 1
// interface positions, z grids (zz, 2); eps (3, xx, yy, zz); stats of
// render_stats_floats written by the forward and read by the backward, which
// needs a workspace of render_workspace_bytes.
int render_max_layers();
size_t render_stats_floats(int ll, int xx, int yy, int zz);
size_t render_workspace_bytes(int ll, int xx, int yy, int zz);
hipError_t render_forward(int ll, int m, int xx, int yy, int zz, int simple, const float* layers,
                          const float* pos, const float* gs, const float* ge, float* eps,
                          float* stats, hipStream_t s);
hipError_t render_backward(int ll, int m, int xx, int yy, int zz, int simple, const float* layers,
                           const float* pos, const float* gs, const float* ge,
                           const float* stats, const float* g_eps, float* g_layers, float* g_pos,
                           char* ws, hipStream_t s);

// lE[window] += g (one snapshot cotangent, (3, ex, ey, ez)).
hipError_t launch_inject(const Geom& g, const pjz_fdtd_problem& p, float* lE, const float* g_out,
                         hipStream_t s);

// grad_eps (3, ex, ey, ez) = sum over the cells each window cell fills of
// (cb / dt) G, which is dL/dca dca/deps + dL/dcb dcb/deps when ca = 1 - s cb.
hipError_t launch_grad_eps(const Geom& g, const pjz_fdtd_problem& p, const float* eps,
                           const float* mask, const char* ws, const Layout& L, float* grad_eps,
                           hipStream_t s);

}  // namespace pjz
