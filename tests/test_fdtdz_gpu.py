"""GPU parity of fdtdz() (through the C ABI) against the FP64 oracle.

Tolerance: relative L2 error <= 1e-4 over all snapshots, FP32 GPU fields vs
the FP64 oracle, for every kernel.  Full-size (512x512x128) runs are checked
against the oracle for 60 steps and, over 200 steps, through properties the
oracle cannot afford there: kernel-vs-kernel agreement, linearity in the
waveform and exact time-shift invariance.  Larger domains (105M and 210M
cells) are checked against the cell-by-cell reference kernel.
"""

import numpy as np
import pytest
import torch

from oracle import c_oracle as C
from oracle import helpers as H
from pjz_amd import fdtdz_torch as F
from pjz_amd import workloads

pytestmark = pytest.mark.gpu

KERNELS = ["reference", "stream", "stream2"]
TOL = 1e-4


def rel_l2(a, b):
  a = np.asarray(a, np.float64)
  b = np.asarray(b, np.float64)
  return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def run_gpu(kw, kernel, **over):
  kw = dict(kw, **over)
  kw["launch_params"] = dict(kernel=kernel)
  out = F.fdtdz(**kw)
  torch.cuda.synchronize()
  return out.cpu().numpy()


def random_problem(axis, X=24, Y=20, Z=32, tt=120, seed=0, widths=(6, 5)):
  rng = np.random.default_rng(seed)
  ex, ey, ez = X - 6, Y - 4, Z - 10
  eps = (1 + 4 * rng.random((3, ex, ey, ez))).astype(np.float32)
  sf = {0: rng.standard_normal((2, 1, Y, Z)), 1: rng.standard_normal((2, X, 1, Z)),
        2: rng.standard_normal((2, 2, X, Y, 1))}[axis].astype(np.float32)
  pos = {0: 8, 1: 6, 2: 13}[axis]
  wave = np.zeros((tt, 2), np.float32)
  t = np.arange(tt)
  wave[:, 0] = np.sin(0.3 * t) * np.exp(-((t - 30) / 12.0) ** 2)
  wave[:, 1] = np.cos(0.3 * t) * np.exp(-((t - 30) / 12.0) ** 2)
  return dict(epsilon=eps, dt=0.5, source_field=sf, source_waveform=wave,
              source_position=pos, absorption_mask=H.absorption_mask(X, Y, 5, 4e-3).astype(
                  np.float32),
              pml_kappa=(1 + rng.random((Z, 2))).astype(np.float32),
              pml_sigma=H.pml_sigma(widths, Z, 0.5, 3.0).astype(np.float32),
              pml_alpha=np.full((Z, 2), 0.05, np.float32), pml_widths=widths,
              output_steps=(tt - 40, tt, 7), use_reduced_precision=False, launch_params=None,
              offset=(3, 2, 5))


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_random_problem_matches_oracle(kernel, axis):
  print(f"case kernel={kernel} axis={axis}", flush=True)
  kw = random_problem(axis, seed=axis)
  ref = C.fdtdz(**kw)
  got = run_gpu(kw, kernel)
  assert got.shape == ref.shape
  assert rel_l2(got, ref) <= TOL


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("shape", [(17, 13, 9), (10, 9, 64), (9, 11, 96), (8, 7, 128),
                                   (6, 5, 200)])
def test_odd_shapes_match_oracle(kernel, shape):
  X, Y, Z = shape
  print(f"case kernel={kernel} shape={shape}", flush=True)
  rng = np.random.default_rng(X * Y)
  kw = random_problem(2, X=20, Y=20, Z=20, tt=50, seed=X * Y, widths=(2, 3))
  kw.update(epsilon=(1 + 4 * rng.random((3, X, Y, Z))).astype(np.float32),
            source_field=rng.standard_normal((2, 2, X, Y, 1)).astype(np.float32),
            source_position=Z // 2, offset=(0, 0, 0),
            absorption_mask=(0.05 * rng.random((3, X, Y))).astype(np.float32),
            pml_kappa=np.ones((Z, 2), np.float32),
            pml_sigma=H.pml_sigma((2, 3), Z, 0.5, 3.0).astype(np.float32),
            pml_alpha=np.full((Z, 2), 0.05, np.float32))
  ref = C.fdtdz(**kw)
  assert rel_l2(run_gpu(kw, kernel), ref) <= TOL


@pytest.mark.parametrize("kernel", KERNELS)
def test_z_as_batch_matches_oracle(kernel):
  kw = random_problem(1, seed=7)
  Z = kw["pml_sigma"].shape[0]
  kw["pml_kappa"] = np.full((Z, 2), np.inf, np.float32)
  kw["pml_sigma"] = np.zeros((Z, 2), np.float32)
  kw["pml_alpha"] = np.zeros((Z, 2), np.float32)
  assert rel_l2(run_gpu(kw, kernel), C.fdtdz(**kw)) <= TOL


@pytest.mark.parametrize("kernel", KERNELS)
def test_layered_64_full_run_matches_oracle(kernel):
  """64x64x64, 2000 steps, snapshots 1990..1999 (the CPU-runnable workload)."""
  kw = workloads.layered_64()
  ref = C.fdtdz(**kw)
  got = run_gpu(kw, kernel)
  err = rel_l2(got, ref)
  print(f"layered_64 {kernel}: rel L2 {err:.3e}, |ref| {np.abs(ref).max():.3e}")
  assert err <= TOL


def test_native_512_short_run_matches_oracle():
  """512x512x128 at its real shape for 60 steps against the oracle."""
  kw = workloads.native_512(steps=60)
  ref = C.fdtdz(**kw)
  for kernel in KERNELS:
    err = rel_l2(run_gpu(kw, kernel), ref)
    print(f"native_512 60 steps {kernel}: rel L2 {err:.3e}")
    assert err <= TOL


def test_native_512_full_size_properties():
  """512x512x128 at its real shape, 200 steps, on the device: linearity
  (doubling the waveform, and superposition of two waveforms) to rounding;
  an even time shift (the two-step kernel keeps its step pairs) bit for bit;
  the one-step kernel equal to the two-step kernel to rounding.

  Doubling is bit-exact only while no value is subnormal (small problems
  show it exactly).  Rounding below 1.2e-38 is absolute, so at the wave's
  leading edge a doubled run rounds differently from the plain one.  From
  that cell on the two runs carry independent ~1-ulp roundings, ~1e-7
  relative over the field.  The FP64 oracle's update evaluated in float32
  on the CPU behaves the same way: the first mismatch is a CPML auxiliary
  value of 7.7e-39 at step 12, and by step 199 the difference is 1.2e-7
  against fields up to 0.46."""
  dev = torch.device("cuda", 0)
  kw = workloads.native_512(steps=200)
  kw["output_steps"] = (194, 200, 4)
  kw = {k: torch.as_tensor(v).to(dev) if isinstance(v, np.ndarray) else v for k, v in kw.items()}
  w = kw["source_waveform"]

  def run(kernel="stream2", **over):
    out = F.fdtdz(**dict(kw, launch_params=dict(kernel=kernel), **over))
    torch.cuda.synchronize()
    return out

  def rel(a, b):
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))

  print("native_512 200 steps: linearity, superposition, time shift, kernels", flush=True)
  base = run()
  assert float(base.abs().max()) > 0
  err = rel(run(source_waveform=2 * w), 2 * base)
  print(f"doubling rel L2 {err:.2e}", flush=True)
  assert err <= 1e-6
  w2 = 0.7 * w.flip(0)
  second = run(source_waveform=w2)
  both = run(source_waveform=w + w2)
  err = rel(both, base + second)
  print(f"superposition rel L2 {err:.2e}", flush=True)
  assert err <= 1e-6
  del second, both
  late = torch.cat([torch.zeros((6, 2), dtype=w.dtype, device=dev), w])
  assert torch.equal(run(source_waveform=late, output_steps=(200, 206, 4)), base)
  err = rel(run("stream"), base)
  print(f"one-step vs two-step rel L2 {err:.2e}", flush=True)
  assert err <= 1e-6


@pytest.mark.parametrize("X,Y", [(1024, 800), (1600, 1024)])
def test_large_domains_match_reference_kernel(X, Y):
  """105M cells (two-step kernel: its store resources hold one level's E, or
  H, within 2^31 bytes up to ~178M cells) and 210M cells (past that: the
  one-step kernel, 64-bit offsets).  A source plane near the high end of x
  puts the non-zero fields at the largest offsets; the default path must
  equal the cell-by-cell reference kernel (independent indexing)."""
  Z, tt = 128, 8
  print(f"large domain {X}x{Y}x{Z}", flush=True)
  rng = np.random.default_rng(21)
  dev = torch.device("cuda", 0)
  eps = torch.ones((3, X, Y, Z), dtype=torch.float32, device=dev)
  sf = torch.as_tensor(rng.standard_normal((2, 1, Y, Z)).astype(np.float32), device=dev)
  wave = torch.as_tensor(np.stack([np.sin(0.3 * np.arange(tt)), np.cos(0.3 * np.arange(tt))],
                                  -1).astype(np.float32), device=dev)
  kw = dict(epsilon=eps, dt=0.5, source_field=sf, source_waveform=wave, source_position=X - 12,
            absorption_mask=torch.zeros((3, X, Y), device=dev),
            pml_kappa=torch.ones((Z, 2), device=dev), pml_sigma=torch.zeros((Z, 2), device=dev),
            pml_alpha=torch.zeros((Z, 2), device=dev), pml_widths=(0, 0),
            output_steps=(tt - 1, tt, 1), use_reduced_precision=False, offset=(0, 0, 0))
  got = F.fdtdz(**dict(kw, launch_params=None))
  ref = F.fdtdz(**dict(kw, launch_params=dict(kernel="reference")))
  torch.cuda.synchronize()
  assert float(torch.linalg.vector_norm(got[:, :, X - 24:])) > 0
  err = float(torch.linalg.vector_norm(got - ref) / torch.linalg.vector_norm(ref))
  print(f"{X * Y * Z / 1e6:.0f}M cells: rel L2 default vs reference kernel {err:.2e}", flush=True)
  assert err <= 1e-6
  del got, ref, eps
  torch.cuda.empty_cache()


def test_time_shift_invariance_is_exact():
  kw = random_problem(0, tt=100, seed=3)
  shift = 18
  w = kw["source_waveform"]
  kw2 = dict(kw, source_waveform=np.concatenate([np.zeros((shift, 2), np.float32), w]),
             output_steps=tuple(s + shift if i < 2 else s
                                for i, s in enumerate(kw["output_steps"])))
  for kernel in KERNELS:
    a = run_gpu(kw, kernel)
    b = run_gpu(kw2, kernel)
    np.testing.assert_array_equal(a, b)
  # The two-step kernel pairs steps differently under an odd shift; its last
  # odd step runs the one-step kernel, so compare to rounding there.
  kw3 = dict(kw, source_waveform=np.concatenate([np.zeros((1, 2), np.float32), w]),
             output_steps=tuple(s + 1 if i < 2 else s for i, s in enumerate(kw["output_steps"])))
  a = run_gpu(kw, "stream2")
  b = run_gpu(kw3, "stream2")
  assert rel_l2(b, a) <= 1e-6


def test_zero_source_gives_zero_fields():
  kw = random_problem(2, seed=4)
  kw["source_waveform"] = np.zeros_like(kw["source_waveform"])
  for kernel in KERNELS:
    assert np.abs(run_gpu(kw, kernel)).max() == 0.0


def test_empty_output_range():
  kw = random_problem(0, seed=5)
  kw["output_steps"] = (50, 50, 1)
  out = F.fdtdz(**kw)
  assert tuple(out.shape) == (0, 3) + kw["epsilon"].shape[1:]


@pytest.mark.parametrize("tile_x", [7, 32])
@pytest.mark.parametrize("kernel,tile_y", [("stream", 2), ("stream", 4), ("stream", 8),
                                           ("stream", 12), ("stream", 15), ("stream2", 4),
                                           ("stream2", 8), ("stream2", 12), ("stream2", 13)])
def test_stream_launch_variants_match_oracle(tile_x, kernel, tile_y):
  kw = random_problem(2, X=40, Y=30, Z=64 + 32, tt=61, seed=11, widths=(8, 8))
  kw["offset"] = (3, 2, 5)
  ref = C.fdtdz(**kw)
  kw["launch_params"] = dict(kernel=kernel, tile_x=tile_x, tile_y=tile_y)
  got = F.fdtdz(**kw).cpu().numpy()
  assert rel_l2(got, ref) <= TOL


@pytest.mark.parametrize("tile_y", [8, 13])
@pytest.mark.parametrize("Z", [40, 64])
def test_one_cell_per_lane_row_tiles_match_oracle(tile_y, Z):
  """zz <= 64 (one cell per lane): the 8- and 13-row two-step tiles."""
  kw = random_problem(1, X=33, Y=29, Z=Z, tt=47, seed=12, widths=(5, 6))
  ref = C.fdtdz(**kw)
  kw["launch_params"] = dict(kernel="stream2", tile_y=tile_y)
  got = F.fdtdz(**kw).cpu().numpy()
  assert rel_l2(got, ref) <= TOL


@pytest.mark.parametrize("Z", [32, 64, 96])
def test_autotune_matches_oracle_and_reuses_choice(Z):
  """The autotuned call (first call times the candidates on the caller's
  workspace, then clears it) gives oracle-exact fields, and the second call
  (cached choice, no tuning) gives the same fields bit for bit."""
  kw = random_problem(1, Z=Z, seed=20 + Z, widths=(4, 6))
  ref = C.fdtdz(**kw)
  first = run_gpu(kw, "autotune")
  second = run_gpu(kw, "autotune")
  assert rel_l2(first, ref) <= TOL
  np.testing.assert_array_equal(first, second)


def test_autotune_with_a_single_step():
  kw = random_problem(0, seed=31)
  kw.update(output_steps=(0, 1, 1))
  assert rel_l2(run_gpu(kw, "autotune"), C.fdtdz(**kw)) <= TOL


def test_copy_probe_copies():
  import ctypes
  from pjz_amd import _lib
  src = torch.arange(1 << 20, dtype=torch.float32, device="cuda")
  dst = torch.zeros_like(src)
  stream = torch.cuda.current_stream()
  _lib.check(_lib.load().pjz_fdtd_copy(ctypes.c_void_p(src.data_ptr()),
                                       ctypes.c_void_p(dst.data_ptr()), src.numel() * 4,
                                       ctypes.c_void_p(stream.cuda_stream)), "copy")
  torch.cuda.synchronize()
  assert torch.equal(src, dst)
