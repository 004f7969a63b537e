/*
 * pjz_fdtd.h -- C ABI of the MI355X (gfx950) Yee FDTD kernel library
 * (libpjzfdtd.so).
 *
 * The library replaces the kernel call pjz makes at
 *   src/pjz/_field.py:254-269   fields = fdtdz_jax.fdtdz(epsilon=..., dt=...,
 *       source_field=..., source_waveform=..., source_position=...,
 *       absorption_mask=..., pml_kappa=..., pml_sigma=..., pml_alpha=...,
 *       pml_widths=..., output_steps=..., use_reduced_precision=...,
 *       launch_params=..., offset=...)
 * whose implementation is the un-vendored PyPI package fdtdz>=1.1.3
 * (setup.py:26).  Every keyword of that call maps onto a field of
 * pjz_fdtd_problem (static, hashable arguments) or onto a device pointer
 * argument (traced arrays).
 *
 * Conventions
 *   - Every array pointer is a DEVICE pointer to contiguous float32, C order,
 *     z fastest, exactly the layouts pjz passes (shapes below).
 *   - The caller owns all memory.  Scratch (field state, coefficients) comes
 *     in as `workspace`, sized by pjz_fdtd_workspace_bytes().  The library
 *     does not allocate device memory.
 *   - Work is enqueued on `stream` (a hipStream_t, NULL = default stream) and
 *     is stream-ordered; no host synchronisation happens inside the calls.
 *   - Return 0 on success.  Non-zero codes: PJZ_FDTD_EINVAL (bad argument,
 *     message names it), PJZ_FDTD_EHIP (HIP runtime error), PJZ_FDTD_ESPACE
 *     (workspace too small).  pjz_fdtd_last_error() returns the calling
 *     thread's last message.
 *   - Calls are re-entrant; concurrent calls must use distinct workspaces.
 */
#ifndef PJZ_FDTD_H_
#define PJZ_FDTD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PJZ_FDTD_ABI_VERSION 1

#if defined(__GNUC__)
#define PJZ_FDTD_API __attribute__((visibility("default")))
#else
#define PJZ_FDTD_API
#endif

#define PJZ_FDTD_OK 0
#define PJZ_FDTD_EINVAL 1
#define PJZ_FDTD_EHIP 2
#define PJZ_FDTD_ESPACE 3

/* Kernel selection (the `launch_params` analogue, field `kernel`). */
#define PJZ_FDTD_KERNEL_AUTO 0
#define PJZ_FDTD_KERNEL_REFERENCE 1 /* two launches per step, one thread/cell */
#define PJZ_FDTD_KERNEL_STREAM 2    /* one fused launch per step, x-sweep    */
#define PJZ_FDTD_KERNEL_STREAM2 3   /* two steps per launch (temporal block) */
#define PJZ_FDTD_KERNEL_AUTOTUNE 4  /* first call per shape times the kernels,   */
                                    /* depths and tilings (synchronising the     */
                                    /* stream), later calls reuse the fastest    */
#define PJZ_FDTD_KERNEL_STREAM3 5   /* three steps per launch (zz <= 64)         */
#define PJZ_FDTD_KERNEL_STREAM4 6   /* four steps per launch (zz <= 64)          */

/* Problem description: the static arguments of fdtdz(...). */
typedef struct pjz_fdtd_problem {
  int32_t xx, yy, zz;    /* domain: absorption_mask is (3,xx,yy); pml_* (zz,2)  */
  int32_t ex, ey, ez;    /* epsilon is (3,ex,ey,ez), placed at (x0,y0,z0)       */
  int32_t x0, y0, z0;    /* `offset`                                             */
  int32_t pml_lo, pml_hi;/* `pml_widths`                                         */
  int32_t src_axis;      /* 0/1/2: source plane normal to x/y/z                  */
  int32_t src_pos;       /* `source_position` (x/y planes: even)                 */
  int32_t src_quad;      /* z plane: 1 for (2,xx,yy,1), 2 for (2,2,xx,yy,1)      */
  int32_t tt;            /* rows of source_waveform (tt,2)                       */
  int32_t out_start, out_stop, out_step; /* `output_steps`                       */
  float dt;              /* `dt`                                                 */
  int32_t flags;         /* bit 0: use_reduced_precision (fp32 storage kept)     */
  int32_t kernel;        /* PJZ_FDTD_KERNEL_*                                    */
  int32_t tile_x, tile_y;/* launch overrides for the stream kernel, 0 = auto     */
                         /* (batches: tile_y 8 or 13 on every zz <= 64 sim    */
                         /* fixes that class's rows per tile)                  */
  int32_t waves_per_simd;/* occupancy hint for the stream kernel, 0 = auto      */
  int32_t ckpt_every;    /* > 0: forward saves a checkpoint every ckpt_every     */
                         /* steps so pjz_fdtd_backward can run on the workspace  */
  int32_t graph_steps;   /* batches: steps per captured, replayed hipGraph    */
                         /* chunk (even, <= 128); 0 = one launch per pair   */
  int32_t ckpt_inner;    /* 0 < ckpt_inner < ckpt_every: two-level checkpoints  */
                         /* (single calls): the forward keeps every ckpt_every-th */
                         /* state; the backward recomputes each such segment     */
                         /* keeping every ckpt_inner-th, then recomputes those   */
                         /* sub-segments capturing E changes (one more forward,  */
                         /* workspace ~ (steps/every + every/inner) states +     */
                         /* inner E changes instead of steps/every + every)     */
} pjz_fdtd_problem;

/* ABI version of the loaded library (PJZ_FDTD_ABI_VERSION). */
PJZ_FDTD_API int pjz_fdtd_abi_version(void);

/* Last error message of the calling thread ("" if none). */
PJZ_FDTD_API const char* pjz_fdtd_last_error(void);

/* Number of snapshots, len(range(*output_steps)). */
PJZ_FDTD_API int pjz_fdtd_num_outputs(const pjz_fdtd_problem* p, int32_t* nn);

/* Checks the problem and reports the workspace size for pjz_fdtd_forward. */
PJZ_FDTD_API int pjz_fdtd_workspace_bytes(const pjz_fdtd_problem* p, size_t* bytes);

/*
 * Forward simulation: replaces fdtdz_jax.fdtdz(...) (src/pjz/_field.py:254-269).
 *   epsilon          (3, ex, ey, ez)
 *   absorption_mask  (3, xx, yy)
 *   pml_kappa/sigma/alpha (zz, 2)
 *   source_field     x: (2,1,yy,zz)  y: (2,xx,1,zz)  z: (src_quad,2,xx,yy,1)
 *   source_waveform  (tt, 2)
 *   out              (nn, 3, ex, ey, ez): E after step s, s in range(*output_steps)
 */
PJZ_FDTD_API int pjz_fdtd_forward(const pjz_fdtd_problem* p, const float* epsilon,
                     const float* absorption_mask, const float* pml_kappa,
                     const float* pml_sigma, const float* pml_alpha,
                     const float* source_field, const float* source_waveform,
                     float* out, void* workspace, size_t workspace_bytes,
                     void* stream);

/*
 * Forward simulation without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  No counterpart in pjz or fdtdz: the
 * library-owned variant of pjz_fdtd_forward (same arguments, results and
 * kernels).  The workspace is a library array of PJZ_FDTD_NOALLOC_MAX_BYTES
 * per GPU, held until the process ends; a problem whose workspace
 * (pjz_fdtd_workspace_bytes) exceeds it returns PJZ_FDTD_ESPACE, and
 * ckpt_every must be 0.  The call allocates nothing and launches only fill,
 * set-up and stepper kernels (an AUTOTUNE kernel id uses a cached choice
 * for the shape, else the default kernel, instead of tuning).  Every call on
 * a GPU uses that GPU's one array: calls that overlap on one GPU (other
 * streams, other threads) race and must be serialised by the caller.
 */
#define PJZ_FDTD_NOALLOC_MAX_BYTES ((size_t)256 << 20)
PJZ_FDTD_API int pjz_fdtd_forward_noalloc(const pjz_fdtd_problem* p, const float* epsilon,
                             const float* absorption_mask, const float* pml_kappa,
                             const float* pml_sigma, const float* pml_alpha,
                             const float* source_field, const float* source_waveform,
                             float* out, void* stream);

/*
 * Batched forward: n independent simulations of uneven size advanced together
 * (persistent two-step launches over a prefix-summed (simulation, tile) work
 * queue), each with the inputs and outputs of pjz_fdtd_forward.  Pointer
 * arguments are HOST arrays of n device pointers.  No counterpart in pjz,
 * which runs one fdtdz() per port in a Python list comprehension
 * (src/pjz/_field.py:354-356); this is the batched form of that loop.
 * Every problem needs zz <= 64 or an even zz <= 128.  With ckpt_every > 0
 * (the same even value in every problem) the call also saves checkpoints for
 * pjz_fdtd_batch_backward in the same workspace.
 */
PJZ_FDTD_API int pjz_fdtd_batch_workspace_bytes(const pjz_fdtd_problem* probs, int n,
                                                size_t* bytes);
PJZ_FDTD_API int pjz_fdtd_batch_forward(const pjz_fdtd_problem* probs, int n,
                           const float* const* epsilon, const float* const* absorption_mask,
                           const float* const* pml_kappa, const float* const* pml_sigma,
                           const float* const* pml_alpha, const float* const* source_field,
                           const float* const* source_waveform, float* const* out,
                           void* workspace, size_t workspace_bytes, void* stream);

/*
 * Batched gradient: pjz_fdtd_backward for every simulation of a preceding
 * pjz_fdtd_batch_forward call with ckpt_every > 0 (same problems, arrays and
 * workspace).  Segments are recomputed by the batched two-step kernel and
 * every reverse step is one launch pair for the whole batch.  grad_out and
 * grad_eps are HOST arrays of n device pointers.  No counterpart in pjz (see
 * pjz_fdtd_backward).
 */
PJZ_FDTD_API int pjz_fdtd_batch_backward(const pjz_fdtd_problem* probs, int n,
                            const float* const* epsilon, const float* const* absorption_mask,
                            const float* const* pml_kappa, const float* const* pml_sigma,
                            const float* const* pml_alpha, const float* const* source_field,
                            const float* const* source_waveform, const float* const* grad_out,
                            float* const* grad_eps, void* workspace, size_t workspace_bytes,
                            void* stream);

/*
 * Gradient of sum_s <grad_out[s], out[s]> with respect to epsilon, by the
 * discrete adjoint of the update (reverse time, transposed curls, CPML and
 * absorber), segment by segment from the checkpoints the preceding
 * pjz_fdtd_forward call saved in the same workspace (p->ckpt_every > 0, same
 * problem and arrays).  No counterpart in pjz, which never differentiates
 * through the kernel (scatter() supplies its own VJP, src/pjz/_field.py:393-398);
 * it is the addition inverse design needs.
 *   grad_out   (nn, 3, ex, ey, ez) cotangent of out
 *   grad_eps   (3, ex, ey, ez), overwritten
 */
PJZ_FDTD_API int pjz_fdtd_backward(const pjz_fdtd_problem* p, const float* epsilon,
                      const float* absorption_mask, const float* pml_kappa,
                      const float* pml_sigma, const float* pml_alpha,
                      const float* source_field, const float* source_waveform,
                      const float* grad_out, float* grad_eps, void* workspace,
                      size_t workspace_bytes, void* stream);

/*
 * Waveguide modes: the subspace iteration of pjz's mode()
 * (src/pjz/_mode.py:99-142, _subspace_iteration, with the operator of
 * _mode.py:22-50), run to convergence inside one stream-ordered launch.
 *   epsilon  (3, xx, yy)          eps_x, eps_y, eps_z of the cross-section in
 *                                 propagate-along-z form
 *   omega    (ww,), shift (ww,)   angular frequency and eigenvalue shift
 *   x        (ww, 2, xx, yy, mm)  start vectors in; on return the last
 *                                 orthonormal basis (modes innermost)
 *   w, err   (ww, mm)             Rayleigh quotients <x_k, A x_k> and residuals
 *                                 |A x_k - w_k x_k|, in basis order (unsorted)
 *   iters    (ww,)                iterations run per frequency
 * Iteration runs until max_iters or until residuals are <= tol.  pjz tests
 * max(err) > tol before each iteration, over all frequencies at once, and
 * large sections do the same (one cooperative launch, several workgroups per
 * frequency).  Small sections (xx * yy * mm < 16384) give each frequency one
 * workgroup, and each frequency stops once its own residuals are <= tol.
 * 1 <= mm <= 10.  Workspace from pjz_fdtd_mode_workspace_bytes.
 */
PJZ_FDTD_API int pjz_fdtd_mode_workspace_bytes(int32_t ww, int32_t xx, int32_t yy, int32_t mm,
                                               size_t* bytes);
PJZ_FDTD_API int pjz_fdtd_mode_solve(int32_t ww, int32_t xx, int32_t yy, int32_t mm,
                                     const float* epsilon, const float* omega, const float* shift,
                                     float* x, float* w, float* err, int32_t* iters,
                                     int32_t max_iters, float tol, void* workspace,
                                     size_t workspace_bytes, void* stream);

/*
 * pjz's density() (src/pjz/_density.py:48-79) for nb design variables of
 * (xx, yy), fp32 device arrays, y fastest: the cone blur of radius `radius`
 * (half-width ceil(radius - 1/2) <= PJZ_FDTD_DENSITY_MAX_HALF_WIDTH) of the
 * edge-padded variable (`_filter`, :17-20), the eta projection (`_project`,
 * :30-31) and the minimum-feature-size loss (`_geom_loss`, :41-45).
 * Forward writes density, loss and the blurred variable `ufilt`, which the
 * backward takes back to compute g_u = dL/du from the cotangents of density
 * and loss (the vector-Jacobian product jax.grad forms through density()),
 * with a workspace of pjz_fdtd_density_workspace_bytes (256-byte aligned).
 */
#define PJZ_FDTD_DENSITY_MAX_HALF_WIDTH 40
PJZ_FDTD_API int pjz_fdtd_density_workspace_bytes(int32_t nb, int32_t xx, int32_t yy,
                                                  float radius, size_t* bytes);
PJZ_FDTD_API int pjz_fdtd_density_forward(int32_t nb, int32_t xx, int32_t yy, float radius,
                                          float alpha, float c, float eta, float eta_lo,
                                          float eta_hi, const float* u, float* density,
                                          float* loss, float* ufilt, void* stream);
PJZ_FDTD_API int pjz_fdtd_density_backward(int32_t nb, int32_t xx, int32_t yy, float radius,
                                           float alpha, float c, float eta, float eta_lo,
                                           float eta_hi, const float* ufilt,
                                           const float* g_density, const float* g_loss,
                                           float* g_u, void* workspace, size_t workspace_bytes,
                                           void* stream);

/*
 * pjz's permittivity render (src/pjz/_epsilon.py:10-108; epsilon() is the
 * case grid_start = k + (-1/2, 0), grid_end = k + (1/2, 1), :152-155): ll <=
 * PJZ_FDTD_RENDER_MAX_LAYERS layered profiles `layers` (ll, 2m xx, 2m yy),
 * ll - 1 increasing interface `positions`, z cell bounds `grid_start`,
 * `grid_end` (zz, 2; column 0 for the x and y components, column 1 for z)
 * -> `eps` (3, xx, yy, zz) by the diagonal sub-pixel average, or the plain
 * volume average when `simple` is non-zero.  fp32 device arrays, last index
 * fastest.  The forward also writes `stats` (12 ll xx yy + 4 ll zz floats:
 * tile moments and the layers' shares of the cells), which the
 * backward takes back with dL/deps to write dL/dlayers and dL/dpositions
 * (the vector-Jacobian product jax.grad forms through the render; none for
 * the z grids), with a workspace of pjz_fdtd_render_workspace_bytes
 * (256-byte aligned).
 */
#define PJZ_FDTD_RENDER_MAX_LAYERS 16
PJZ_FDTD_API int pjz_fdtd_render_workspace_bytes(int32_t ll, int32_t xx, int32_t yy, int32_t zz,
                                                 size_t* bytes);
PJZ_FDTD_API int pjz_fdtd_render_forward(int32_t ll, int32_t m, int32_t xx, int32_t yy,
                                         int32_t zz, int32_t simple, const float* layers,
                                         const float* positions, const float* grid_start,
                                         const float* grid_end, float* eps, float* stats,
                                         void* stream);
PJZ_FDTD_API int pjz_fdtd_render_backward(int32_t ll, int32_t m, int32_t xx, int32_t yy,
                                          int32_t zz, int32_t simple, const float* layers,
                                          const float* positions, const float* grid_start,
                                          const float* grid_end, const float* stats,
                                          const float* g_eps, float* g_layers,
                                          float* g_positions, void* workspace,
                                          size_t workspace_bytes, void* stream);

/*
 * Diagnostic: the stepper (PJZ_FDTD_KERNEL_REFERENCE .. STREAM4) a forward
 * call with this problem runs on the current device: AUTO and AUTOTUNE
 * resolved (AUTOTUNE to its cached choice, or AUTO before the first call).
 * No counterpart in pjz: bench.py reports it beside the kernel's timings.
 */
PJZ_FDTD_API int pjz_fdtd_stepper(const pjz_fdtd_problem* p, int32_t* kernel);

/*
 * Diagnostic: copies `bytes` (a multiple of 16; both pointers 16-byte
 * aligned) from src to dst on the device with nontemporal loads and stores,
 * stream-ordered.  No counterpart in pjz: bench.py times it on 1 GiB to
 * report the device-copy bandwidth the steppers' measured HBM traffic is
 * compared against.
 */
PJZ_FDTD_API int pjz_fdtd_copy(const void* src, void* dst, size_t bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PJZ_FDTD_H_ */
