// C ABI entry points of libpjzfdtd.so: argument checking, workspace layout,
// set-up and dispatch to the steppers.  Declarations: include/pjz_fdtd.h.
#include <stdio.h>
#include <string.h>
#include <stdarg.h>
#include <limits.h>

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
  if (p->kernel < PJZ_FDTD_KERNEL_AUTO || p->kernel > PJZ_FDTD_KERNEL_STREAM)
    return fail(PJZ_FDTD_EINVAL, "unknown kernel id %d", p->kernel);
  if (p->kernel == PJZ_FDTD_KERNEL_STREAM && p->zz > 256)
    return fail(PJZ_FDTD_EINVAL, "the stream kernel needs zz <= 256, got %d", p->zz);
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

}  // namespace

extern "C" {

int pjz_fdtd_abi_version(void) { return PJZ_FDTD_ABI_VERSION; }

const char* pjz_fdtd_last_error(void) { return g_err; }

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
  *bytes = pjz::make_layout(geom_of(*p)).total;
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
  pjz::Layout L = pjz::make_layout(g);
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
  pjz::OutWin win{p->x0, p->y0, p->z0, p->ex, p->ey, p->ez, out};
  int nsteps = p->out_start + (nn - 1) * p->out_step + 1;
  int kernel = p->kernel;
  if (kernel == PJZ_FDTD_KERNEL_AUTO)
    kernel = g.Z <= 256 ? PJZ_FDTD_KERNEL_STREAM : PJZ_FDTD_KERNEL_REFERENCE;
  if (kernel == PJZ_FDTD_KERNEL_STREAM)
    e = pjz::run_stream(g, *p, absorption_mask, src, win, ws, L, nsteps, s);
  else
    e = pjz::run_reference(g, *p, absorption_mask, src, win, ws, L, nsteps, s);
  if (e != hipSuccess) return hip_fail(e, "step launch");
  return PJZ_FDTD_OK;
}

}  // extern "C"
