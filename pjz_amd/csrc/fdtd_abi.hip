// C ABI entry points of libpjzfdtd.so: argument checking, workspace layout,
// set-up and dispatch to the steppers.  Declarations: include/pjz_fdtd.h.
#include <stdio.h>
#include <string.h>
#include <stdarg.h>
#include <limits.h>
#include <stdint.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "fdtd_common.h"

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int hip_fail(hipError_t e, const char* where) {
  return fail(PJZ_FDTD_EHIP, "%s: HIP error %d (%s)", where, (int)e, hipGetErrorString(e));
}

int count_outputs(const pjz_fdtd_problem& p) {
  if (p.out_stop <= p.out_start) return 0;
  return (p.out_stop - p.out_start + p.out_step - 1) / p.out_step;
}

pjz::Geom geom_of(const pjz_fdtd_problem& p);

// Checks every static argument; mirrors the shape contract pjz relies on
// (src/pjz/_field.py:222-269).
int check(const pjz_fdtd_problem* p) {
  if (!p) return fail(PJZ_FDTD_EINVAL, "problem is NULL");
  if (p->xx < 1 || p->yy < 1 || p->zz < 1)
    return fail(PJZ_FDTD_EINVAL, "domain (%d,%d,%d) must be positive", p->xx, p->yy, p->zz);
  long long cells = (long long)p->xx * p->yy * p->zz;
  if (3 * cells >= (long long)INT_MAX)
    return fail(PJZ_FDTD_EINVAL, "domain of %lld cells exceeds 2^31/3 cells", cells);
  if (p->ex < 1 || p->ey < 1 || p->ez < 1)
    return fail(PJZ_FDTD_EINVAL, "epsilon extent (%d,%d,%d) must be positive", p->ex, p->ey,
                p->ez);
  if (p->x0 < 0 || p->y0 < 0 || p->z0 < 0 || p->x0 + p->ex > p->xx || p->y0 + p->ey > p->yy ||
      p->z0 + p->ez > p->zz)
    return fail(PJZ_FDTD_EINVAL,
                "epsilon (3,%d,%d,%d) at offset (%d,%d,%d) does not fit the (%d,%d,%d) domain",
                p->ex, p->ey, p->ez, p->x0, p->y0, p->z0, p->xx, p->yy, p->zz);
  if (p->pml_lo < 0 || p->pml_hi < 0 || p->pml_lo + p->pml_hi > p->zz)
    return fail(PJZ_FDTD_EINVAL, "pml_widths (%d,%d) do not fit zz=%d", p->pml_lo, p->pml_hi,
                p->zz);
  if (!(p->dt > 0.0f)) return fail(PJZ_FDTD_EINVAL, "dt must be positive");
  if (p->src_axis < 0 || p->src_axis > 2)
    return fail(PJZ_FDTD_EINVAL, "source axis %d is not 0, 1 or 2", p->src_axis);
  int dims[3] = {p->xx, p->yy, p->zz};
  if (p->src_pos < 0 || p->src_pos >= dims[p->src_axis])
    return fail(PJZ_FDTD_EINVAL, "source_position %d outside [0, %d)", p->src_pos,
                dims[p->src_axis]);
  if (p->src_axis < 2 && (p->src_pos % 2) != 0)
    return fail(PJZ_FDTD_EINVAL, "source_position %d must be even for an x/y source plane",
                p->src_pos);
  if (p->src_axis == 2 && (p->src_quad < 1 || p->src_quad > 2))
    return fail(PJZ_FDTD_EINVAL, "z source needs 1 or 2 quadratures, got %d", p->src_quad);
  if (p->tt < 1) return fail(PJZ_FDTD_EINVAL, "source_waveform needs at least one row");
  if (p->out_step < 1) return fail(PJZ_FDTD_EINVAL, "output_steps step must be >= 1");
  if (p->out_start < 0) return fail(PJZ_FDTD_EINVAL, "output_steps start must be >= 0");
  int nn = count_outputs(*p);
  if (nn > 0) {
    int last = p->out_start + (nn - 1) * p->out_step;
    if (last >= p->tt)
      return fail(PJZ_FDTD_EINVAL, "last output step %d needs waveform rows > %d (tt=%d)", last,
                  last, p->tt);
  }
  if (p->kernel < PJZ_FDTD_KERNEL_AUTO || p->kernel > PJZ_FDTD_KERNEL_STREAM4)
    return fail(PJZ_FDTD_EINVAL, "unknown kernel id %d", p->kernel);
  if (p->kernel == PJZ_FDTD_KERNEL_STREAM && p->zz > 256)
    return fail(PJZ_FDTD_EINVAL, "the stream kernel needs zz <= 256, got %d", p->zz);
  if (p->kernel == PJZ_FDTD_KERNEL_STREAM2 && !pjz::stream2_supported(geom_of(*p)))
    return fail(PJZ_FDTD_EINVAL,
                "the two-step stream kernel needs zz <= 64, an even zz <= 128 or a zz <= 256 "
                "divisible by 4, got %d", p->zz);
  if (p->kernel == PJZ_FDTD_KERNEL_STREAM3 && !pjz::stream_deep_supported(geom_of(*p), 3))
    return fail(PJZ_FDTD_EINVAL, "the three-step stream kernel needs zz <= 64, got %d", p->zz);
  if (p->kernel == PJZ_FDTD_KERNEL_STREAM4 && !pjz::stream_deep_supported(geom_of(*p), 4))
    return fail(PJZ_FDTD_EINVAL, "the four-step stream kernel needs zz <= 64, got %d", p->zz);
  if (p->ckpt_every < 0 || p->ckpt_inner < 0)
    return fail(PJZ_FDTD_EINVAL, "ckpt_every and ckpt_inner must be >= 0, got %d, %d",
                p->ckpt_every, p->ckpt_inner);
  if (p->ckpt_inner > 0 && p->ckpt_every <= 0)
    return fail(PJZ_FDTD_EINVAL, "ckpt_inner %d needs ckpt_every > 0", p->ckpt_inner);
  return PJZ_FDTD_OK;
}

pjz::Geom geom_of(const pjz_fdtd_problem& p) {
  pjz::Geom g;
  g.X = p.xx;
  g.Y = p.yy;
  g.Z = p.zz;
  g.lo = p.pml_lo;
  g.hi = p.pml_hi;
  g.NP = p.pml_lo + p.pml_hi;
  g.N = (long long)p.xx * p.yy * p.zz;
  return g;
}

// ---------------------------------------------------------------- autotuning
// Kernel and tiling measured fastest for a problem shape on a device; the
// cache is process-wide and guarded by a mutex.
struct Tuned {
  int kernel, tile_x, tile_y;
};
using TuneKey = std::tuple<int, int, int, int, int, int>;  // device, xx, yy, zz, lo, hi
std::mutex g_tune_mu;
std::map<TuneKey, Tuned> g_tuned;

TuneKey tune_key(const pjz_fdtd_problem& p) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  return TuneKey{dev, p.xx, p.yy, p.zz, p.pml_lo, p.pml_hi};
}

bool tuned_choice(const pjz_fdtd_problem& p, Tuned* t) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  auto it = g_tuned.find(tune_key(p));
  if (it == g_tuned.end()) return false;
  *t = it->second;
  return true;
}

hipError_t run_steps(const pjz::Geom& g, const pjz_fdtd_problem& p, const float* mask,
                     const pjz::Src& src, const pjz::OutWin& win, char* ws,
                     const pjz::Layout& L, int b, int end, int* cur, float* capture,
                     hipStream_t s);

// ------------------------------------------------- library-held scratch
// Workspace of pjz_fdtd_forward_noalloc: one array per GPU, held until the
// process ends.  In host code `scratch` only names the symbol; its device
// address on the current GPU comes from hipGetSymbolAddress, looked up once
// per device and kept in g_scratch_ptr.
constexpr size_t kScratchBytes = PJZ_FDTD_NOALLOC_MAX_BYTES;
__device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current GPU's scratch array (nullptr and a HIP error
// on failure).
hipError_t scratch_address(void** ptr) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = p;
  }
  *ptr = g_scratch_ptr[dev];
  return hipSuccess;
}

hipError_t clear_state(const pjz::Geom& g, const pjz::Layout& L, char* ws, hipStream_t s) {
  hipError_t e = pjz::launch_zero((float*)(ws + L.E0), (L.E1 - L.E0) / sizeof(float), s);
  if (e == hipSuccess && g.NP > 0)
    e = pjz::launch_zero((float*)(ws + L.psi), (L.zt - L.psi) / sizeof(float), s);
  return e;
}

// Times the candidate kernels and tilings on a few steps of this problem
// (from a cleared state, no snapshots, on the caller's workspace and
// stream), caches the fastest for the shape and leaves the state cleared.
// Synchronises the stream: only the first call for a shape pays for it.
hipError_t autotune(const pjz::Geom& g, const pjz_fdtd_problem& p, const float* mask,
                    const pjz::Src& src, char* ws, const pjz::Layout& L, hipStream_t s) {
  std::vector<Tuned> cands;
  const bool cpl2 = g.Z > 64 && g.Z <= 128 && g.Z % 2 == 0;
  if (pjz::stream2_supported(g)) {
    cands.push_back({PJZ_FDTD_KERNEL_STREAM2, 0, 0});
    cands.push_back({PJZ_FDTD_KERNEL_STREAM2, 64, 0});
    if (cpl2) cands.push_back({PJZ_FDTD_KERNEL_STREAM2, 0, 4});
    // At zz <= 64 the default (tile_y 0) picks 13 or 8 rows by the CU count;
    // both are timed explicitly.
    if (g.Z <= 64) {
      cands.push_back({PJZ_FDTD_KERNEL_STREAM2, 0, 13});
      cands.push_back({PJZ_FDTD_KERNEL_STREAM2, 0, 8});
    }
  }
  // Deeper temporal blocking: three or four steps per pass.
  if (pjz::stream_deep_supported(g, 3)) cands.push_back({PJZ_FDTD_KERNEL_STREAM3, 0, 0});
  if (pjz::stream_deep_supported(g, 4)) cands.push_back({PJZ_FDTD_KERNEL_STREAM4, 0, 0});
  if (g.Z <= 256) {
    cands.push_back({PJZ_FDTD_KERNEL_STREAM, 0, 0});
    if (cpl2) cands.push_back({PJZ_FDTD_KERNEL_STREAM, 0, 4});
  }
  cands.push_back({PJZ_FDTD_KERNEL_REFERENCE, 0, 0});
  // Up to 12 steps (a whole number of launches at every depth), within the
  // waveform's rows (the run's own step count).
  const int nsteps = std::min(12, pjz::steps_of(p)) & ~1;
  if (nsteps < 2) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tuned[tune_key(p)] = cands[0];
    return hipSuccess;
  }
  pjz::OutWin nowin{p.x0, p.y0, p.z0, p.ex, p.ey, p.ez, nullptr};
  hipEvent_t t0, t1;
  hipError_t e = hipEventCreate(&t0);
  if (e != hipSuccess) return e;
  e = hipEventCreate(&t1);
  if (e != hipSuccess) {
    (void)hipEventDestroy(t0);
    return e;
  }
  Tuned best = cands[0];
  float best_ms = 1e30f;
  for (size_t i = 0; i < cands.size() && e == hipSuccess; ++i) {
    pjz_fdtd_problem q = p;
    q.kernel = cands[i].kernel;
    q.tile_x = cands[i].tile_x;
    q.tile_y = cands[i].tile_y;
    // One untimed pass loads the code, then the timed one.
    for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {
      int cur = 0;
      e = clear_state(g, L, ws, s);
      if (e == hipSuccess && rep == 1) e = hipEventRecord(t0, s);
      if (e == hipSuccess) e = run_steps(g, q, mask, src, nowin, ws, L, 0, nsteps, &cur, nullptr, s);
      if (e == hipSuccess && rep == 1) e = hipEventRecord(t1, s);
    }
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventSynchronize(t1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
    if (e == hipSuccess && ms < best_ms) {
      best_ms = ms;
      best = cands[i];
    }
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (e == hipSuccess) e = clear_state(g, L, ws, s);
  if (e == hipSuccess) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tuned[tune_key(p)] = best;
  }
  return e;
}

// The stepper AUTO runs: three steps per launch on large grids of zz <= 64,
// else the two-step kernel where supported, else the one-step kernel if
// zz <= 256, else the reference kernel.
int auto_kernel(const pjz::Geom& g) {
  if (pjz::stream_deep_preferred(g)) return PJZ_FDTD_KERNEL_STREAM3;
  if (pjz::stream2_supported(g)) return PJZ_FDTD_KERNEL_STREAM2;
  return g.Z <= 256 ? PJZ_FDTD_KERNEL_STREAM : PJZ_FDTD_KERNEL_REFERENCE;
}

// Steps [b, end) with the kernel the problem selects (AUTO: auto_kernel;
// AUTOTUNE: the cached choice for the shape, else AUTO).
hipError_t run_steps(const pjz::Geom& g, const pjz_fdtd_problem& p, const float* mask,
                     const pjz::Src& src, const pjz::OutWin& win, char* ws,
                     const pjz::Layout& L, int b, int end, int* cur, float* capture,
                     hipStream_t s) {
  int kernel = p.kernel;
  Tuned t;
  if (kernel == PJZ_FDTD_KERNEL_AUTOTUNE) {
    if (tuned_choice(p, &t)) {
      pjz_fdtd_problem q = p;
      q.kernel = t.kernel;
      q.tile_x = t.tile_x;
      q.tile_y = t.tile_y;
      return run_steps(g, q, mask, src, win, ws, L, b, end, cur, capture, s);
    }
    kernel = PJZ_FDTD_KERNEL_AUTO;
  }
  if (kernel == PJZ_FDTD_KERNEL_AUTO) kernel = auto_kernel(g);
  pjz::StepRange r{b, end, cur, capture};
  if (kernel == PJZ_FDTD_KERNEL_STREAM2)
    return pjz::run_stream2(g, p, mask, src, win, ws, L, r, s);
  if (kernel == PJZ_FDTD_KERNEL_STREAM3)
    return pjz::run_stream_deep(g, p, mask, src, win, ws, L, r, 3, s);
  if (kernel == PJZ_FDTD_KERNEL_STREAM4)
    return pjz::run_stream_deep(g, p, mask, src, win, ws, L, r, 4, s);
  if (kernel == PJZ_FDTD_KERNEL_STREAM) return pjz::run_stream(g, p, mask, src, win, ws, L, r, s);
  return pjz::run_reference(g, p, mask, src, win, ws, L, r, s);
}

// Checkpoint k holds E, H, the psi_H pair and the psi_E pair of the state that
// lives in ping-pong set `cur`.
struct StateParts {
  float* E;
  float* H;
  float* psiH;
  float* psiE;
};

StateParts state_parts(const pjz::Geom& g, const pjz::Layout& L, char* ws, int cur) {
  float* psi = (float*)(ws + L.psi);
  const long long P = (long long)g.X * g.Y * g.NP;
  return {(float*)(ws + (cur ? L.E1 : L.E0)), (float*)(ws + (cur ? L.H1 : L.H0)),
          psi + (cur ? pjz::PSI_SET : 0) * P, psi + ((cur ? pjz::PSI_SET : 0) + pjz::PSI_EX) * P};
}

hipError_t copy_state(const pjz::Geom& g, const pjz::Layout& L, char* ws, int cur, int k,
                      bool save, hipStream_t s) {
  StateParts st = state_parts(g, L, ws, cur);
  float* c = (float*)(ws + L.ckpt + pjz::ckpt_stride(L) * (size_t)k);
  const size_t n3 = 3 * (size_t)g.N, p2 = 2 * (size_t)g.X * g.Y * (size_t)g.NP;
  float* parts[4] = {st.E, st.H, st.psiH, st.psiE};
  const size_t sizes[4] = {n3, n3, p2, p2};
  size_t off = 0;
  for (int i = 0; i < 4; ++i) {
    if (sizes[i] == 0) continue;
    hipError_t e = save ? pjz::launch_copy(c + off, parts[i], sizes[i], s)
                        : pjz::launch_copy(parts[i], c + off, sizes[i], s);
    if (e != hipSuccess) return e;
    off += sizes[i];
  }
  return hipSuccess;
}

hipError_t save_state(const pjz::Geom& g, const pjz::Layout& L, char* ws, int cur, int k,
                      hipStream_t s) {
  return copy_state(g, L, ws, cur, k, true, s);
}

// Restores checkpoint k into ping-pong set 0.
hipError_t restore_state(const pjz::Geom& g, const pjz::Layout& L, char* ws, int k,
                         hipStream_t s) {
  return copy_state(g, L, ws, 0, k, false, s);
}

}  // namespace

extern "C" {

int pjz_fdtd_abi_version(void) { return PJZ_FDTD_ABI_VERSION; }

const char* pjz_fdtd_last_error(void) { return g_err; }

static int check_mode(int32_t ww, int32_t xx, int32_t yy, int32_t mm) {
  if (ww < 1 || ww > 65535) return fail(PJZ_FDTD_EINVAL, "mode: ww = %d out of [1, 65535]", ww);
  if (xx < 1 || yy < 1) return fail(PJZ_FDTD_EINVAL, "mode: cross-section %d x %d", xx, yy);
  if (mm < 1 || mm > pjz::kModeMaxModes)
    return fail(PJZ_FDTD_EINVAL, "mode: num_modes = %d out of [1, %d]", mm, pjz::kModeMaxModes);
  if (2LL * xx * yy * mm > (long long)INT_MAX / 2)
    return fail(PJZ_FDTD_EINVAL, "mode: 2 x %d x %d x %d values exceed the kernel's 32-bit indexing",
                xx, yy, mm);
  return PJZ_FDTD_OK;
}

int pjz_fdtd_mode_workspace_bytes(int32_t ww, int32_t xx, int32_t yy, int32_t mm, size_t* bytes) {
  int rc = check_mode(ww, xx, yy, mm);
  if (rc) return rc;
  if (!bytes) return fail(PJZ_FDTD_EINVAL, "mode: bytes is NULL");
  *bytes = pjz::mode_workspace_floats(ww, xx, yy, mm) * sizeof(float);
  return PJZ_FDTD_OK;
}

int pjz_fdtd_mode_solve(int32_t ww, int32_t xx, int32_t yy, int32_t mm, const float* epsilon,
                        const float* omega, const float* shift, float* x, float* w, float* err,
                        int32_t* iters, int32_t max_iters, float tol, void* workspace,
                        size_t workspace_bytes, void* stream) {
  int rc = check_mode(ww, xx, yy, mm);
  if (rc) return rc;
  if (!epsilon || !omega || !shift || !x || !w || !err || !iters || !workspace)
    return fail(PJZ_FDTD_EINVAL, "mode: NULL array");
  if (max_iters < 0) return fail(PJZ_FDTD_EINVAL, "mode: max_iters = %d < 0", max_iters);
  if (!(tol >= 0.0f)) return fail(PJZ_FDTD_EINVAL, "mode: tol must be >= 0");
  const size_t need = pjz::mode_workspace_floats(ww, xx, yy, mm) * sizeof(float);
  if (workspace_bytes < need)
    return fail(PJZ_FDTD_ESPACE, "mode: workspace of %zu bytes, need %zu", workspace_bytes, need);
  hipError_t e = pjz::launch_mode(ww, xx, yy, mm, epsilon, omega, shift, x, w, err, iters,
                                  max_iters, tol, (float*)workspace, (hipStream_t)stream);
  return e == hipSuccess ? PJZ_FDTD_OK : hip_fail(e, "mode");
}

int pjz_fdtd_copy(const void* src, void* dst, size_t bytes, void* stream) {
  if (!src || !dst) return fail(PJZ_FDTD_EINVAL, "copy: NULL pointer");
  if (bytes % 16 || (uintptr_t)src % 16 || (uintptr_t)dst % 16)
    return fail(PJZ_FDTD_EINVAL, "copy: bytes and both pointers must be multiples of 16");
  if (bytes == 0) return PJZ_FDTD_OK;
  hipError_t e = pjz::launch_copy_stream(dst, src, bytes / 16, (hipStream_t)stream);
  return e == hipSuccess ? PJZ_FDTD_OK : hip_fail(e, "copy");
}

namespace {
// Arguments of the density entry points.
int check_density(int32_t nb, int32_t xx, int32_t yy, float radius, float eta_lo,
                  float eta_hi) {
  if (nb < 1 || xx < 1 || yy < 1)
    return fail(PJZ_FDTD_EINVAL, "density: nb, xx, yy must be >= 1, got %d, %d, %d", nb, xx, yy);
  if ((long long)nb * xx * yy > (1LL << 31) / 4)
    return fail(PJZ_FDTD_EINVAL, "density: nb * xx * yy too large");
  if (!(radius > 0.0f) || pjz::density_half_width(radius) > PJZ_FDTD_DENSITY_MAX_HALF_WIDTH)
    return fail(PJZ_FDTD_EINVAL, "density: radius must be in (0, %d.5), got %g",
                PJZ_FDTD_DENSITY_MAX_HALF_WIDTH, (double)radius);
  if (!(eta_lo <= eta_hi))
    return fail(PJZ_FDTD_EINVAL, "density: eta_lo must be <= eta_hi");
  return PJZ_FDTD_OK;
}
}  // namespace

int pjz_fdtd_density_workspace_bytes(int32_t nb, int32_t xx, int32_t yy, float radius,
                                     size_t* bytes) {
  int rc = check_density(nb, xx, yy, radius, 0.0f, 0.0f);
  if (rc) return rc;
  if (!bytes) return fail(PJZ_FDTD_EINVAL, "bytes is NULL");
  *bytes = pjz::density_workspace_bytes(nb, xx, yy, radius);
  return PJZ_FDTD_OK;
}

int pjz_fdtd_density_forward(int32_t nb, int32_t xx, int32_t yy, float radius, float alpha,
                             float c, float eta, float eta_lo, float eta_hi, const float* u,
                             float* density, float* loss, float* ufilt, void* stream) {
  int rc = check_density(nb, xx, yy, radius, eta_lo, eta_hi);
  if (rc) return rc;
  if (!u || !density || !loss || !ufilt)
    return fail(PJZ_FDTD_EINVAL, "density: a required pointer argument is NULL");
  hipError_t e = pjz::density_forward(nb, xx, yy, radius, alpha, c, eta, eta_lo, eta_hi, u,
                                      density, loss, ufilt, (hipStream_t)stream);
  return e == hipSuccess ? PJZ_FDTD_OK : hip_fail(e, "density forward");
}

int pjz_fdtd_density_backward(int32_t nb, int32_t xx, int32_t yy, float radius, float alpha,
                              float c, float eta, float eta_lo, float eta_hi, const float* ufilt,
                              const float* g_density, const float* g_loss, float* g_u,
                              void* workspace, size_t workspace_bytes, void* stream) {
  int rc = check_density(nb, xx, yy, radius, eta_lo, eta_hi);
  if (rc) return rc;
  if (!ufilt || !g_density || !g_loss || !g_u || !workspace)
    return fail(PJZ_FDTD_EINVAL, "density: a required pointer argument is NULL");
  const size_t need = pjz::density_workspace_bytes(nb, xx, yy, radius);
  if (workspace_bytes < need)
    return fail(PJZ_FDTD_ESPACE, "density: workspace of %zu bytes, %zu needed", workspace_bytes,
                need);
  if (((uintptr_t)workspace & 255) != 0)
    return fail(PJZ_FDTD_EINVAL, "density: workspace must be 256-byte aligned");
  hipError_t e = pjz::density_backward(nb, xx, yy, radius, alpha, c, eta, eta_lo, eta_hi, ufilt,
                                       g_density, g_loss, g_u, (char*)workspace,
                                       (hipStream_t)stream);
  return e == hipSuccess ? PJZ_FDTD_OK : hip_fail(e, "density backward");
}

namespace {
int check_render(int32_t ll, int32_t m, int32_t xx, int32_t yy, int32_t zz) {
  if (ll < 1 || ll > PJZ_FDTD_RENDER_MAX_LAYERS)
    return fail(PJZ_FDTD_EINVAL, "render: 1 <= layers <= %d, got %d", PJZ_FDTD_RENDER_MAX_LAYERS,
                ll);
  if (m < 1 || xx < 1 || yy < 1 || zz < 1)
    return fail(PJZ_FDTD_EINVAL, "render: m, xx, yy, zz must be >= 1, got %d, %d, %d, %d", m, xx,
                yy, zz);
  if ((long long)ll * 4 * m * m * xx * yy >= (1LL << 31) / 4 ||
      3LL * xx * yy * zz >= (1LL << 31) / 4)
    return fail(PJZ_FDTD_EINVAL, "render: arrays too large");
  return PJZ_FDTD_OK;
}
}  // namespace

int pjz_fdtd_render_workspace_bytes(int32_t ll, int32_t xx, int32_t yy, int32_t zz,
                                    size_t* bytes) {
  int rc = check_render(ll, 1, xx, yy, zz);
  if (rc) return rc;
  if (!bytes) return fail(PJZ_FDTD_EINVAL, "bytes is NULL");
  *bytes = pjz::render_workspace_bytes(ll, xx, yy, zz);
  return PJZ_FDTD_OK;
}

int pjz_fdtd_render_forward(int32_t ll, int32_t m, int32_t xx, int32_t yy, int32_t zz,
                            int32_t simple, const float* layers, const float* positions,
                            const float* grid_start, const float* grid_end, float* eps,
                            float* stats, void* stream) {
  int rc = check_render(ll, m, xx, yy, zz);
  if (rc) return rc;
  if (!layers || (ll > 1 && !positions) || !grid_start || !grid_end || !eps || !stats)
    return fail(PJZ_FDTD_EINVAL, "render: a required pointer argument is NULL");
  hipError_t e = pjz::render_forward(ll, m, xx, yy, zz, simple != 0, layers, positions,
                                     grid_start, grid_end, eps, stats, (hipStream_t)stream);
  return e == hipSuccess ? PJZ_FDTD_OK : hip_fail(e, "render forward");
}

int pjz_fdtd_render_backward(int32_t ll, int32_t m, int32_t xx, int32_t yy, int32_t zz,
                             int32_t simple, const float* layers, const float* positions,
                             const float* grid_start, const float* grid_end, const float* stats,
                             const float* g_eps, float* g_layers, float* g_positions,
                             void* workspace, size_t workspace_bytes, void* stream) {
  int rc = check_render(ll, m, xx, yy, zz);
  if (rc) return rc;
  if (!layers || (ll > 1 && (!positions || !g_positions)) || !grid_start || !grid_end ||
      !stats || !g_eps || !g_layers || !workspace)
    return fail(PJZ_FDTD_EINVAL, "render: a required pointer argument is NULL");
  const size_t need = pjz::render_workspace_bytes(ll, xx, yy, zz);
  if (workspace_bytes < need)
    return fail(PJZ_FDTD_ESPACE, "render: workspace of %zu bytes, %zu needed", workspace_bytes,
                need);
  if (((uintptr_t)workspace & 255) != 0)
    return fail(PJZ_FDTD_EINVAL, "render: workspace must be 256-byte aligned");
  hipError_t e = pjz::render_backward(ll, m, xx, yy, zz, simple != 0, layers, positions,
                                      grid_start, grid_end, stats, g_eps, g_layers, g_positions,
                                      (char*)workspace, (hipStream_t)stream);
  return e == hipSuccess ? PJZ_FDTD_OK : hip_fail(e, "render backward");
}

int pjz_fdtd_stepper(const pjz_fdtd_problem* p, int32_t* kernel) {
  int rc = check(p);
  if (rc) return rc;
  if (!kernel) return fail(PJZ_FDTD_EINVAL, "kernel is NULL");
  const pjz::Geom g = geom_of(*p);
  int k = p->kernel;
  Tuned t;
  if (k == PJZ_FDTD_KERNEL_AUTOTUNE) k = tuned_choice(*p, &t) ? t.kernel : PJZ_FDTD_KERNEL_AUTO;
  *kernel = k == PJZ_FDTD_KERNEL_AUTO ? auto_kernel(g) : k;
  return PJZ_FDTD_OK;
}

int pjz_fdtd_num_outputs(const pjz_fdtd_problem* p, int32_t* nn) {
  int rc = check(p);
  if (rc) return rc;
  if (!nn) return fail(PJZ_FDTD_EINVAL, "nn is NULL");
  *nn = count_outputs(*p);
  return PJZ_FDTD_OK;
}

int pjz_fdtd_workspace_bytes(const pjz_fdtd_problem* p, size_t* bytes) {
  int rc = check(p);
  if (rc) return rc;
  if (!bytes) return fail(PJZ_FDTD_EINVAL, "bytes is NULL");
  *bytes = pjz::make_layout(geom_of(*p), *p).total;
  return PJZ_FDTD_OK;
}

int pjz_fdtd_forward(const pjz_fdtd_problem* p, const float* epsilon,
                     const float* absorption_mask, const float* pml_kappa,
                     const float* pml_sigma, const float* pml_alpha,
                     const float* source_field, const float* source_waveform, float* out,
                     void* workspace, size_t workspace_bytes, void* stream) {
  int rc = check(p);
  if (rc) return rc;
  int nn = count_outputs(*p);
  if (nn == 0) return PJZ_FDTD_OK;
  if (!epsilon || !absorption_mask || !pml_kappa || !pml_sigma || !pml_alpha || !source_field ||
      !source_waveform || !out || !workspace)
    return fail(PJZ_FDTD_EINVAL, "a required pointer argument is NULL");
  pjz::Geom g = geom_of(*p);
  pjz::Layout L = pjz::make_layout(g, *p);
  if (workspace_bytes < L.total)
    return fail(PJZ_FDTD_ESPACE, "workspace of %zu bytes, %zu needed", workspace_bytes, L.total);
  if (((uintptr_t)workspace & 255) != 0)
    return fail(PJZ_FDTD_EINVAL, "workspace must be 256-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  // Clear E0, H0 (contiguous) and psi; build cb and the z table.
  hipError_t e = pjz::launch_prepare(g, *p, epsilon, absorption_mask, pml_kappa, pml_sigma,
                                     pml_alpha, (float*)(ws + L.cb), (float*)(ws + L.zt), s);
  if (e != hipSuccess) return hip_fail(e, "prepare");
  e = pjz::launch_zero((float*)(ws + L.E0), (L.E1 - L.E0) / sizeof(float), s);
  if (e == hipSuccess && g.NP > 0)
    e = pjz::launch_zero((float*)(ws + L.psi), (L.zt - L.psi) / sizeof(float), s);
  if (e != hipSuccess) return hip_fail(e, "clear state");
  pjz::Src src{p->src_axis, p->src_pos, p->src_quad, source_field, source_waveform};
  Tuned tuned;
  if (p->kernel == PJZ_FDTD_KERNEL_AUTOTUNE && !tuned_choice(*p, &tuned)) {
    e = autotune(g, *p, absorption_mask, src, ws, L, s);
    if (e != hipSuccess) return hip_fail(e, "autotune");
  }
  pjz::OutWin win{p->x0, p->y0, p->z0, p->ex, p->ey, p->ez, out};
  const int nsteps = pjz::steps_of(*p);
  int cur = 0;
  if (L.nck > 0) {
    // Checkpoint the state before steps 0, S, 2S, ... for pjz_fdtd_backward.
    for (int k = 0; k < L.nck && e == hipSuccess; ++k) {
      const int b = k * L.S1, end = b + L.S1 < nsteps ? b + L.S1 : nsteps;
      e = save_state(g, L, ws, cur, k, s);
      if (e == hipSuccess) e = run_steps(g, *p, absorption_mask, src, win, ws, L, b, end, &cur,
                                         nullptr, s);
    }
  } else {
    e = run_steps(g, *p, absorption_mask, src, win, ws, L, 0, nsteps, &cur, nullptr, s);
  }
  if (e != hipSuccess) return hip_fail(e, "step launch");
  return PJZ_FDTD_OK;
}

int pjz_fdtd_forward_noalloc(const pjz_fdtd_problem* p, const float* epsilon,
                             const float* absorption_mask, const float* pml_kappa,
                             const float* pml_sigma, const float* pml_alpha,
                             const float* source_field, const float* source_waveform,
                             float* out, void* stream) {
  int rc = check(p);
  if (rc) return rc;
  if (p->ckpt_every > 0)
    return fail(PJZ_FDTD_EINVAL,
                "forward_noalloc runs without checkpoints: ckpt_every must be 0, got %d",
                p->ckpt_every);
  const size_t need = pjz::make_layout(geom_of(*p), *p).total;
  if (need > kScratchBytes)
    return fail(PJZ_FDTD_ESPACE,
                "forward_noalloc: a (%d,%d,%d) domain needs %zu bytes of workspace, above the "
                "library's %zu-byte scratch array (use pjz_fdtd_forward with a workspace)",
                p->xx, p->yy, p->zz, need, kScratchBytes);
  void* ws = nullptr;
  hipError_t e = scratch_address(&ws);
  if (e != hipSuccess) return hip_fail(e, "forward_noalloc: scratch address");
  // Autotuning would synchronise and time candidates; use a cached choice
  // for the shape if there is one, else the default kernel.
  pjz_fdtd_problem q = *p;
  Tuned t;
  if (q.kernel == PJZ_FDTD_KERNEL_AUTOTUNE && !tuned_choice(q, &t)) q.kernel = PJZ_FDTD_KERNEL_AUTO;
  return pjz_fdtd_forward(&q, epsilon, absorption_mask, pml_kappa, pml_sigma, pml_alpha,
                          source_field, source_waveform, out, ws, kScratchBytes, stream);
}

}  // extern "C"

namespace {

int max_steps(const pjz_fdtd_problem* probs, int n) {
  int m = 0;
  for (int i = 0; i < n; ++i) m = std::max(m, pjz::steps_of(probs[i]));
  return m;
}

// Workspace plan of a batch: one prepared single-simulation layout per item
// (with checkpoints and adjoint buffers when ckpt_every > 0), then the batch
// tables.
struct BatchPlanHost {
  std::vector<pjz::Layout> Ls;
  std::vector<size_t> offs;
  size_t meta_off = 0, adj_off = 0, adj2_off = 0, total = 0;
  int S = 0;  // common checkpoint spacing, 0 = forward only
};

int batch_plan(const pjz_fdtd_problem* probs, int n, BatchPlanHost& bp) {
  if (!probs || n < 1) return fail(PJZ_FDTD_EINVAL, "a batch needs at least one problem");
  bp.Ls.resize(n);
  bp.offs.resize(n);
  bp.S = probs[0].ckpt_every > 0 ? probs[0].ckpt_every : 0;
  size_t off = 0;
  for (int i = 0; i < n; ++i) {
    int rc = check(&probs[i]);
    if (rc) {
      char msg[sizeof(g_err)];
      snprintf(msg, sizeof(msg), "%s", g_err);
      return fail(rc, "batch item %d: %s", i, msg);
    }
    const int S = probs[i].ckpt_every > 0 ? probs[i].ckpt_every : 0;
    if (S != bp.S)
      return fail(PJZ_FDTD_EINVAL, "batch item %d: ckpt_every %d differs from item 0's %d", i, S,
                  bp.S);
    if (S % 2 != 0)
      return fail(PJZ_FDTD_EINVAL, "batch item %d: ckpt_every must be even, got %d", i, S);
    if (probs[i].ckpt_inner != 0)
      return fail(PJZ_FDTD_EINVAL, "batch item %d: batches keep one level of checkpoints "
                  "(ckpt_inner must be 0, got %d)", i, probs[i].ckpt_inner);
    pjz::Geom g = geom_of(probs[i]);
    if (!pjz::batch2_supported(g))
      return fail(PJZ_FDTD_EINVAL, "batch item %d: zz must be <= 64 or even and <= 128, got %d", i,
                  g.Z);
    bp.Ls[i] = pjz::make_layout(g, probs[i]);
    bp.offs[i] = off;
    off += pjz::align256(bp.Ls[i].total);
  }
  bp.meta_off = off;
  std::vector<pjz::BatchItem> items(n);
  for (int i = 0; i < n; ++i)
    items[i] = {geom_of(probs[i]), &probs[i], nullptr, {}, {}, nullptr, &bp.Ls[i], nullptr};
  bp.adj_off = off + pjz::align256(pjz::batch2_meta_bytes(items.data(), n));
  bp.adj2_off = bp.adj_off + (bp.S > 0 ? pjz::align256(pjz::adjoint_batch_meta_bytes(n, bp.S)) : 0);
  bp.total = bp.adj2_off + (bp.S > 0 ? pjz::adjoint2_batch_meta_bytes(n, bp.S) : 0);
  return PJZ_FDTD_OK;
}

// Shared checks of the batched entry points; on success `items` holds the
// prepared simulations (coefficients built; state cleared, or adjoint
// buffers cleared when `adjoint`).
int batch_setup(const pjz_fdtd_problem* probs, int n, const BatchPlanHost& bp,
                const float* const* epsilon, const float* const* absorption_mask,
                const float* const* pml_kappa, const float* const* pml_sigma,
                const float* const* pml_alpha, const float* const* source_field,
                const float* const* source_waveform, float* const* out, void* workspace,
                size_t workspace_bytes, bool adjoint, hipStream_t s,
                std::vector<pjz::BatchItem>& items) {
  if (!epsilon || !absorption_mask || !pml_kappa || !pml_sigma || !pml_alpha || !source_field ||
      !source_waveform || !out || !workspace)
    return fail(PJZ_FDTD_EINVAL, "a required pointer array is NULL");
  if (workspace_bytes < bp.total)
    return fail(PJZ_FDTD_ESPACE, "workspace of %zu bytes, %zu needed", workspace_bytes, bp.total);
  if (((uintptr_t)workspace & 255) != 0)
    return fail(PJZ_FDTD_EINVAL, "workspace must be 256-byte aligned");
  char* base = (char*)workspace;
  items.resize(n);
  for (int i = 0; i < n; ++i) {
    if (!epsilon[i] || !absorption_mask[i] || !pml_kappa[i] || !pml_sigma[i] || !pml_alpha[i] ||
        !source_field[i] || !source_waveform[i] || !out[i])
      return fail(PJZ_FDTD_EINVAL, "batch item %d: a required pointer is NULL", i);
    const pjz_fdtd_problem& q = probs[i];
    pjz::Geom g = geom_of(q);
    char* ws = base + bp.offs[i];
    const pjz::Layout& L = bp.Ls[i];
    hipError_t e = pjz::launch_prepare(g, q, epsilon[i], absorption_mask[i], pml_kappa[i],
                                       pml_sigma[i], pml_alpha[i], (float*)(ws + L.cb),
                                       (float*)(ws + L.zt), s);
    if (adjoint) {
      // Clear both adjoint sets and G (contiguous from L.lE_[0] to the end).
      if (e == hipSuccess && L.total > L.lE_[0])
        e = pjz::launch_zero((float*)(ws + L.lE_[0]), (L.total - L.lE_[0]) / 4, s);
    } else {
      if (e == hipSuccess) e = pjz::launch_zero((float*)(ws + L.E0), (L.E1 - L.E0) / 4, s);
      if (e == hipSuccess && g.NP > 0)
        e = pjz::launch_zero((float*)(ws + L.psi), (L.zt - L.psi) / 4, s);
    }
    if (e != hipSuccess) return hip_fail(e, "batch set-up");
    items[i] = {g, &probs[i], absorption_mask[i],
                pjz::Src{q.src_axis, q.src_pos, q.src_quad, source_field[i], source_waveform[i]},
                pjz::OutWin{q.x0, q.y0, q.z0, q.ex, q.ey, q.ez, adjoint ? nullptr : out[i]}, ws,
                &bp.Ls[i], adjoint && L.nck > 0 ? (float*)(ws + L.seg) : nullptr};
  }
  return PJZ_FDTD_OK;
}

}  // namespace

extern "C" {

int pjz_fdtd_batch_workspace_bytes(const pjz_fdtd_problem* probs, int n, size_t* bytes) {
  BatchPlanHost bp;
  int rc = batch_plan(probs, n, bp);
  if (rc) return rc;
  if (!bytes) return fail(PJZ_FDTD_EINVAL, "bytes is NULL");
  *bytes = bp.total;
  return PJZ_FDTD_OK;
}

int pjz_fdtd_batch_forward(const pjz_fdtd_problem* probs, int n, const float* const* epsilon,
                           const float* const* absorption_mask, const float* const* pml_kappa,
                           const float* const* pml_sigma, const float* const* pml_alpha,
                           const float* const* source_field,
                           const float* const* source_waveform, float* const* out,
                           void* workspace, size_t workspace_bytes, void* stream) {
  BatchPlanHost bp;
  int rc = batch_plan(probs, n, bp);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  std::vector<pjz::BatchItem> items;
  rc = batch_setup(probs, n, bp, epsilon, absorption_mask, pml_kappa, pml_sigma, pml_alpha,
                   source_field, source_waveform, out, workspace, workspace_bytes, false, s, items);
  if (rc) return rc;
  char* base = (char*)workspace;
  pjz::Batch2Plan plan;
  hipError_t e = pjz::batch2_prepare(items.data(), n, base + bp.meta_off, plan, s, bp.S == 0);
  if (e == hipSuccess && bp.S == 0) e = pjz::batch2_forward(plan, s);
  // With checkpoints: before the pairs of segment k, save the state of step
  // k*S (buffer set (k*S/2) & 1) of every simulation that reaches it.
  for (int k = 0; bp.S > 0 && e == hipSuccess && k * bp.S < 2 * plan.npairs; ++k) {
    const int b = k * bp.S;
    for (int i = 0; i < n && e == hipSuccess; ++i)
      if (k < bp.Ls[i].nck)
        e = copy_state(items[i].g, bp.Ls[i], items[i].ws, (b / 2) & 1, k, true, s);
    if (e == hipSuccess) e = pjz::batch2_pairs(plan, b / 2, (b + bp.S) / 2, s);
  }
  const hipError_t r = pjz::batch2_release(plan, s);
  if (e == hipSuccess) e = r;
  if (e != hipSuccess) return hip_fail(e, "batch launch");
  return PJZ_FDTD_OK;
}

int pjz_fdtd_batch_backward(const pjz_fdtd_problem* probs, int n, const float* const* epsilon,
                            const float* const* absorption_mask, const float* const* pml_kappa,
                            const float* const* pml_sigma, const float* const* pml_alpha,
                            const float* const* source_field,
                            const float* const* source_waveform, const float* const* grad_out,
                            float* const* grad_eps, void* workspace, size_t workspace_bytes,
                            void* stream) {
  BatchPlanHost bp;
  int rc = batch_plan(probs, n, bp);
  if (rc) return rc;
  if (bp.S <= 0)
    return fail(PJZ_FDTD_EINVAL,
                "batch backward needs ckpt_every > 0 (and a batch forward call with it)");
  if (!grad_out || !grad_eps) return fail(PJZ_FDTD_EINVAL, "a required pointer array is NULL");
  for (int i = 0; i < n; ++i)
    if (!grad_out[i] || !grad_eps[i])
      return fail(PJZ_FDTD_EINVAL, "batch item %d: a required pointer is NULL", i);
  hipStream_t s = (hipStream_t)stream;
  std::vector<pjz::BatchItem> items;
  rc = batch_setup(probs, n, bp, epsilon, absorption_mask, pml_kappa, pml_sigma, pml_alpha,
                   source_field, source_waveform, grad_eps, workspace, workspace_bytes, true, s,
                   items);
  if (rc) return rc;
  char* base = (char*)workspace;
  const int last_seg = (std::max(1, max_steps(probs, n)) - 1) / bp.S;
  auto seg_end = [&](int k) {
    int end = k * bp.S;
    for (int i = 0; i < n; ++i)
      if (k < bp.Ls[i].nck) end = std::max(end, std::min((k + 1) * bp.S, pjz::steps_of(probs[i])));
    return end;
  };
  auto is_out = [&](int i, int m) {
    const pjz_fdtd_problem& q = probs[i];
    return m < pjz::steps_of(q) && m >= q.out_start && m < q.out_stop &&
           (m - q.out_start) % q.out_step == 0;
  };
  // The reverse schedule, fixed up front: launches of one or two reversed
  // steps in decreasing step order, segment by segment.  A pair (m, m-1) is
  // used unless a snapshot cotangent enters some simulation at m-1 or some
  // simulation's first reversed step is m-1.  lstart[i]: the launch a
  // simulation joins at (its adjoint set is (launch - lstart) & 1).
  struct RevLaunch {
    int m, b;
    bool two;
  };
  std::vector<RevLaunch> sched;
  std::vector<int> lstart(n, 0);
  std::vector<bool> joined(n, false);
  for (int k = last_seg; k >= 0; --k) {
    const int b = k * bp.S;
    for (int m = seg_end(k) - 1; m >= b;) {
      bool two = m - 1 >= b;
      for (int i = 0; i < n && two; ++i)
        if (pjz::steps_of(probs[i]) == m || is_out(i, m - 1)) two = false;
      for (int i = 0; i < n; ++i)
        if (!joined[i] && m < pjz::steps_of(probs[i])) {
          joined[i] = true;
          lstart[i] = (int)sched.size();
        }
      sched.push_back({m, b, two});
      m -= two ? 2 : 1;
    }
  }
  pjz::Batch2Plan plan;
  pjz::AdjBatchPlan aplan;
  pjz::Adj2BatchPlan a2plan;
  hipError_t e = pjz::batch2_prepare(items.data(), n, base + bp.meta_off, plan, s);
  if (e == hipSuccess)
    e = pjz::adjoint_batch_prepare(items.data(), n, bp.S, lstart.data(), base + bp.adj_off, aplan,
                                   s);
  if (e == hipSuccess)
    e = pjz::adjoint2_batch_prepare(items.data(), n, bp.S, lstart.data(), base + bp.adj2_off,
                                    a2plan, s);
  size_t L = 0;
  for (int k = last_seg; k >= 0 && e == hipSuccess; --k) {
    const int b = k * bp.S;
    // Recompute segment k of every simulation that reaches it from its
    // checkpoint, capturing D^n = E^n - E^(n+1) (slot n - b).
    for (int i = 0; i < n && e == hipSuccess; ++i) {
      const pjz::Layout& Li = bp.Ls[i];
      if (k >= Li.nck) continue;
      e = copy_state(items[i].g, Li, items[i].ws, (b / 2) & 1, k, false, s);
    }
    if (e == hipSuccess) e = pjz::batch2_capture_from(plan, b, s);
    if (e == hipSuccess) e = pjz::batch2_pairs(plan, b / 2, (seg_end(k) + 1) / 2, s);
    if (e == hipSuccess) e = pjz::adjoint_batch_begin_segment(aplan, s);
    if (e == hipSuccess) e = pjz::adjoint2_batch_begin_segment(a2plan, s);
    for (; L < sched.size() && sched[L].b == b && e == hipSuccess; ++L) {
      const int m = sched[L].m;
      for (int i = 0; i < n && e == hipSuccess; ++i) {
        if (!is_out(i, m)) continue;
        const pjz_fdtd_problem& q = probs[i];
        const long long gstride = 3LL * q.ex * q.ey * q.ez;
        e = pjz::launch_inject(items[i].g, q,
                               (float*)(items[i].ws + bp.Ls[i].lE_[((int)L - lstart[i]) & 1]),
                               grad_out[i] + gstride * ((m - q.out_start) / q.out_step), s);
      }
      if (e != hipSuccess) break;
      e = sched[L].two ? pjz::launch_adjoint2_batch_step(a2plan, (int)L, m, b, s)
                       : pjz::launch_adjoint_batch_step(aplan, (int)L, m, b, s);
    }
  }
  for (int i = 0; i < n && e == hipSuccess; ++i)
    e = bp.Ls[i].nck == 0  // no steps: nothing depends on epsilon
            ? pjz::launch_zero(grad_eps[i], 3 * (size_t)probs[i].ex * probs[i].ey * probs[i].ez, s)
            : pjz::launch_grad_eps(items[i].g, probs[i], epsilon[i], absorption_mask[i],
                                   items[i].ws, bp.Ls[i], grad_eps[i], s);
  const hipError_t r = pjz::batch2_release(plan, s);
  if (e == hipSuccess) e = r;
  if (e != hipSuccess) return hip_fail(e, "batch backward launch");
  return PJZ_FDTD_OK;
}

int pjz_fdtd_backward(const pjz_fdtd_problem* p, const float* epsilon,
                      const float* absorption_mask, const float* pml_kappa,
                      const float* pml_sigma, const float* pml_alpha,
                      const float* source_field, const float* source_waveform,
                      const float* grad_out, float* grad_eps, void* workspace,
                      size_t workspace_bytes, void* stream) {
  int rc = check(p);
  if (rc) return rc;
  if (p->ckpt_every <= 0)
    return fail(PJZ_FDTD_EINVAL, "backward needs ckpt_every > 0 (and a forward call with it)");
  if (!epsilon || !absorption_mask || !pml_kappa || !pml_sigma || !pml_alpha || !source_field ||
      !source_waveform || !grad_out || !grad_eps || !workspace)
    return fail(PJZ_FDTD_EINVAL, "a required pointer argument is NULL");
  pjz::Geom g = geom_of(*p);
  pjz::Layout L = pjz::make_layout(g, *p);
  if (workspace_bytes < L.total)
    return fail(PJZ_FDTD_ESPACE, "workspace of %zu bytes, %zu needed", workspace_bytes, L.total);
  if (((uintptr_t)workspace & 255) != 0)
    return fail(PJZ_FDTD_EINVAL, "workspace must be 256-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  const long long f3 = 3 * g.N;
  const size_t P = (size_t)g.X * g.Y * (size_t)(g.NP > 0 ? g.NP : 1);
  const int nsteps = pjz::steps_of(*p);
  hipError_t e = hipSuccess;
  if (nsteps == 0) {
    e = pjz::launch_zero(grad_eps, 3 * (size_t)p->ex * p->ey * p->ez, s);
    return e == hipSuccess ? PJZ_FDTD_OK : hip_fail(e, "zero grad");
  }
  e = pjz::launch_prepare(g, *p, epsilon, absorption_mask, pml_kappa, pml_sigma, pml_alpha,
                          (float*)(ws + L.cb), (float*)(ws + L.zt), s);
  // Clear both adjoint sets and G (contiguous from L.lE_[0] to the end).
  if (e == hipSuccess)
    e = pjz::launch_zero((float*)(ws + L.lE_[0]), (L.total - L.lE_[0]) / 4, s);
  if (e != hipSuccess) return hip_fail(e, "backward set-up");
  // The reference kernel choice also selects the cell-by-cell reverse step.
  const bool stream_adj =
      p->kernel != PJZ_FDTD_KERNEL_REFERENCE && pjz::adjoint_stream_supported(g);
  // Two reversed steps per launch unless the one-step kernels were asked for.
  const bool stream2_adj = stream_adj && p->kernel != PJZ_FDTD_KERNEL_STREAM &&
                           pjz::adjoint_stream2_supported(g);
  pjz::Src src{p->src_axis, p->src_pos, p->src_quad, source_field, source_waveform};
  pjz::OutWin nowin{p->x0, p->y0, p->z0, p->ex, p->ey, p->ez, nullptr};
  float* seg = (float*)(ws + L.seg);
  const long long gstride = 3LL * p->ex * p->ey * p->ez;
  int acur = 0;  // adjoint set holding the current cotangents
  (void)P;
  auto is_out = [&](int n) {
    return n >= p->out_start && n < p->out_stop && (n - p->out_start) % p->out_step == 0;
  };
  // Steps [b, end) from the state in checkpoint slot k: recompute them
  // capturing D^n = E^n - E^(n+1), then reverse them.
  auto reverse_segment = [&](int k, int b, int end) {
    int cur = 0;
    hipError_t e = restore_state(g, L, ws, k, s);
    if (e == hipSuccess)
      e = run_steps(g, *p, absorption_mask, src, nowin, ws, L, b, end, &cur, seg, s);
    for (int n = end - 1; n >= b && e == hipSuccess;) {
      if (is_out(n))
        e = pjz::launch_inject(g, *p, (float*)(ws + L.lE_[acur]),
                               grad_out + gstride * ((n - p->out_start) / p->out_step), s);
      if (e != hipSuccess) break;
      if (stream2_adj && n - 1 >= b && !is_out(n - 1)) {
        // Steps n and n-1 in one launch (no cotangent enters between them).
        e = pjz::launch_adjoint_stream2(g, *p, absorption_mask, ws, L, seg + (n - b) * f3,
                                        seg + (n - 1 - b) * f3, acur, s);
        acur = 1 - acur;
        n -= 2;
      } else if (stream_adj) {
        e = pjz::launch_adjoint_stream(g, *p, absorption_mask, ws, L, seg + (n - b) * f3, acur,
                                       s);
        acur = 1 - acur;
        --n;
      } else {
        e = pjz::launch_adjoint_step(g, *p, absorption_mask, ws, L, seg + (n - b) * f3, acur, s);
        --n;
      }
    }
    return e;
  };
  for (int k = L.nck - 1; k >= 0 && e == hipSuccess; --k) {
    const int b1 = k * L.S1, e1 = b1 + L.S1 < nsteps ? b1 + L.S1 : nsteps;
    if (L.nfine == 0) {
      e = reverse_segment(k, b1, e1);
      continue;
    }
    // Two levels: recompute the S1 segment from its checkpoint keeping the
    // state every S steps (slots nck + j), then reverse its sub-segments.
    int cur = 0;
    e = restore_state(g, L, ws, k, s);
    int nsub = 0;
    for (int b = b1; b < e1 && e == hipSuccess; b += L.S, ++nsub) {
      e = save_state(g, L, ws, cur, L.nck + nsub, s);
      const int end = b + L.S < e1 ? b + L.S : e1;
      // The last sub-segment's steps are recomputed again below.
      if (e == hipSuccess && end < e1)
        e = run_steps(g, *p, absorption_mask, src, nowin, ws, L, b, end, &cur, nullptr, s);
    }
    for (int j = nsub - 1; j >= 0 && e == hipSuccess; --j) {
      const int b = b1 + j * L.S, end = b + L.S < e1 ? b + L.S : e1;
      e = reverse_segment(L.nck + j, b, end);
    }
  }
  if (e == hipSuccess)
    e = pjz::launch_grad_eps(g, *p, epsilon, absorption_mask, ws, L, grad_eps, s);
  if (e != hipSuccess) return hip_fail(e, "backward launch");
  return PJZ_FDTD_OK;
}

}  // extern "C"
