"""Benchmark of the fdtdz() hot path on one MI355X.

One bench "step" is one full fdtdz() simulation of the workload (every FDTD
time step of it), inputs already resident in HBM.  Default workload:
native_512 -- 512x512x128 cells, 10000 FDTD steps, 8 snapshots, fp32.

Prints ONE JSON line: Gcell-updates/s over the timed simulations, the HBM
roofline of the stepper kernel (algorithmic bytes per launch / measured
average launch duration, HIP events on the kernel's stream), and the FP64
CPU oracle timed on a bounded sample of the same workload on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                       [--kernel auto|reference|stream|stream2|stream3|stream4]
                       [--no-cpu-baseline]

--workload native_512_z64 runs the same at 512x512x64, where the default
kernel takes three steps per launch; --workload native_512_z96 at
512x512x96, pjz's default height.

Other workloads (one JSON line each): --workload grad_256 (forward +
adjoint gradient, 256x256x128, 6000 steps), --workload batch64 (64 uneven
simulations batched against 64 sequential calls) and --workload
design_batch (64 simulations of 128x128x64, forward + gradient, batched
against sequential calls) and --workload design_batch_384 (the same at the
full 384x384x128, 8000 steps, in memory-bounded waves), --workload mode (the
mode solver kernel), --workload density and --workload render (pjz's
density() and epsilon() on their HIP kernels, forward + gradient).
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
  sys.path.insert(0, ROOT)

METRIC = ("Gcell-updates/s + % HBM roofline; batched fwd+grad vs plain at "
          "≤1e-4 rel err")
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD published


def algorithmic_bytes_per_cell(zz, pml_lo, pml_hi):
  """fp32 compulsory HBM bytes per cell per pass of one time step.

  read E, H (24 B) + read cb (12 B) + write E, H (24 B) = 60 B; CPML psi
  (4 components, read + write) 32 B on the pml_lo + pml_hi PML cells of each
  z column; absorber mask 12 B per (x, y) column.
  """
  return 60.0 + 32.0 * (pml_lo + pml_hi) / zz + 12.0 / zz


def cpu_baseline(kw, target_s=10.0):
  """Times the FP64 C oracle (OpenMP) on a bounded sample of the workload."""
  from oracle import c_oracle
  from pjz_amd import workloads
  try:
    threads = len(os.sched_getaffinity(0))
  except AttributeError:
    threads = os.cpu_count() or 1
  threads = max(1, min(threads, 16))
  cells = workloads.cells(kw)

  def run(n):
    t0 = time.perf_counter()
    c_oracle.fdtdz(**kw, threads=threads, nsteps=n)
    return time.perf_counter() - t0

  # Set-up (coefficients, allocation) is timed at 0 steps and subtracted.
  t_setup = run(0)
  t_probe = run(10) - t_setup
  n = int(max(10, min(4000, 10 * target_s / max(t_probe, 1e-3))))
  wall = run(n) - t_setup
  # Host roofline: the stepper's FP64 algorithmic bytes per cell-update (one
  # step per pass: twice the fp32 figure) against a multi-threaded 1 GiB host
  # copy on the same cores.
  lo, hi = kw["pml_widths"]
  bpc64 = 2.0 * algorithmic_bytes_per_cell(kw["pml_sigma"].shape[0], lo, hi)
  copy_gbs = host_copy_gbs(threads)
  achieved = cells * n * bpc64 / wall / 1e9
  full = workloads.time_steps(kw)
  return {"value": round(cells * n / wall / 1e9, 5), "unit": "Gcell-updates/s",
          "cores": threads, "kind": "port", "extrapolated": True,
          "sample": f"FP64 C oracle (oracle/fdtd_ref.c, OpenMP) on the full "
                    f"{kw['epsilon'].shape[1]}x{kw['epsilon'].shape[2]}x"
                    f"{kw['epsilon'].shape[3]} grid, {n} of the run's {full} steps in "
                    f"{wall:.1f} s (set-up of {t_setup:.1f} s excluded); the rate per "
                    f"cell-update is extrapolated to the whole run "
                    f"({cells * full / (cells * n / wall) / 60:.0f} min on these cores)",
          "algorithmic_gbs": round(achieved, 2), "host_copy_gbs": round(copy_gbs, 2),
          "roofline_frac": round(achieved / copy_gbs, 4)}


def host_copy_gbs(threads, gib=1, reps=3):
  """Host DRAM bandwidth (GB/s, read + write): best of `reps` torch copies of
  `gib` GiB on `threads` threads."""
  import torch
  torch.set_num_threads(threads)
  a = torch.ones((gib << 30) // 4, dtype=torch.float32)
  b = torch.empty_like(a)
  b.copy_(a)
  best = 1e30
  for _ in range(reps):
    t0 = time.perf_counter()
    b.copy_(a)
    best = min(best, time.perf_counter() - t0)
  return 2.0 * (gib << 30) / best / 1e9


def copy_ceiling_gbs(dev, gib=1, reps=10):
  """Device-copy bandwidth (GB/s, read + write) of the library's nontemporal
  copy kernel on `gib` GiB, HIP events on the launching stream."""
  import ctypes
  import torch
  from pjz_amd import _lib
  lib = _lib.load()
  n = gib << 30
  src = torch.ones(n // 4, dtype=torch.float32, device=dev)
  dst = torch.empty_like(src)
  stream = torch.cuda.current_stream(dev)
  call = lambda: _lib.check(lib.pjz_fdtd_copy(ctypes.c_void_p(src.data_ptr()),
                                              ctypes.c_void_p(dst.data_ptr()), n,
                                              ctypes.c_void_p(stream.cuda_stream)), "copy")
  call(), call()
  start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
  start.record(stream)
  for _ in range(reps):
    call()
  end.record(stream)
  torch.cuda.synchronize()
  if not torch.equal(src[:1024], dst[:1024]):
    raise RuntimeError("copy kernel produced wrong data")
  return 2.0 * n * reps / (start.elapsed_time(end) * 1e-3) / 1e9


def measured_traffic(kernel, workload):
  """HBM bytes per launch from the committed rocprofv3 counter summary."""
  path = os.path.join(ROOT, "profiles", f"k_{kernel}_{workload}.json")
  if not os.path.exists(path):
    return None
  with open(path) as f:
    return json.load(f)["traffic_bytes_per_launch"]


def bench_grad(args):
  """Forward + adjoint gradient of the overlap objective on grad_256."""
  import torch
  from pjz_amd import fdtdz, workloads
  steps = args.time_steps or 6000
  kw, mode, omegas = workloads.grad_256(steps=steps)
  dev = torch.device("cuda", 0)
  eps0 = torch.as_tensor(kw.pop("epsilon"), device=dev)
  for k, v in list(kw.items()):
    if hasattr(v, "shape") and not isinstance(v, torch.Tensor):
      kw[k] = torch.as_tensor(v, device=dev)
  kw["launch_params"] = dict(kernel=args.kernel, ckpt_every=256)
  cells = workloads.cells(kw)
  tsteps = workloads.time_steps(kw)

  def run(with_grad):
    e = eps0.clone().requires_grad_(with_grad)
    out = fdtdz(e, **kw)
    loss = workloads.overlap_objective(out, mode, omegas, kw["output_steps"])
    if with_grad:
      loss.backward()
    return loss

  for _ in range(args.warmup):
    run(True)
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  for _ in range(args.steps):
    run(False)
  torch.cuda.synchronize()
  fwd = (time.perf_counter() - t0) / args.steps
  t0 = time.perf_counter()
  for _ in range(args.steps):
    run(True)
  torch.cuda.synchronize()
  both = (time.perf_counter() - t0) / args.steps
  line = {"metric": METRIC, "value": round(cells * tsteps / both / 1e9, 3),
          "unit": "Gcell-updates/s (forward steps per second of forward + gradient)",
          "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
          "ms_per_step": round(both * 1e3, 3), "higher_is_better": True, "scaling": "weak",
          "vs_baseline": None, "dtype": "f32", "data": "synthetic",
          "config": {"workload": f"grad_256: 256x256x128 cells, {tsteps} steps, objective "
                                 f"sum_w |<E(w) at x=220, mode>|^2, dL/deps by the adjoint, "
                                 f"checkpoint every 256 steps",
                     "cells": int(cells), "time_steps": int(tsteps),
                     "forward_s": round(fwd, 4), "forward_plus_gradient_s": round(both, 4),
                     "gradient_over_forward": round(both / fwd, 3)}}
  print(json.dumps(line), flush=True)


def bench_batch(args):
  """64 uneven simulations: one batched call against 64 sequential calls."""
  import torch
  from pjz_amd import fdtdz, fdtdz_batch, workloads
  steps = args.time_steps or 4000
  calls = workloads.batch64(steps=steps)
  cells = sum(int(c["epsilon"].shape[1] * c["epsilon"].shape[2] * c["epsilon"].shape[3])
              for c in calls)
  for _ in range(args.warmup):
    fdtdz_batch(calls[:4])
    for c in calls[:4]:
      fdtdz(**c)
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  for _ in range(args.steps):
    fdtdz_batch(calls)
  torch.cuda.synchronize()
  tb = (time.perf_counter() - t0) / args.steps
  t0 = time.perf_counter()
  for _ in range(args.steps):
    for c in calls:
      fdtdz(**c)
  torch.cuda.synchronize()
  ts = (time.perf_counter() - t0) / args.steps
  line = {"metric": METRIC, "value": round(cells * steps / tb / 1e9, 3),
          "unit": "Gcell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
          "ms_per_step": round(tb * 1e3, 3), "higher_is_better": True, "scaling": "weak",
          "vs_baseline": None, "dtype": "f32", "data": "synthetic",
          "config": {"workload": f"batch64: 64 seeded uneven simulations (xx, yy in "
                                 f"128..512, zz in 64, 128), {steps} steps, 4 snapshots",
                     "cells": cells, "time_steps": steps, "batched_s": round(tb, 4),
                     "sequential_s": round(ts, 4), "speedup_vs_sequential": round(ts / tb, 3)}}
  print(json.dumps(line), flush=True)


def bench_design(args):
  """Batched inverse-design iteration: 16 variants x 2 ports x 2 frequencies
  (64 simulations), forward + adjoint gradient of the summed overlap
  objective, as one batch against 64 sequential single-simulation calls with
  three kernels: the default (two-step) path, the one-step streaming kernel
  (one launch per step) and the cell-by-cell reference kernel (two launches
  per step, one thread per cell)."""
  import torch
  from pjz_amd import fdtdz, fdtdz_batch, workloads
  steps = args.time_steps or 2000
  calls, variant_of, monitors = workloads.design_batch(steps=steps)
  nv = max(variant_of) + 1
  base_eps = [None] * nv
  for c, v in zip(calls, variant_of):
    base_eps[v] = c["epsilon"]
  cells = sum(int(c["epsilon"].shape[1] * c["epsilon"].shape[2] * c["epsilon"].shape[3])
              for c in calls)

  def objective(i, out):
    w, mon_x, mode = monitors[i]
    return workloads.overlap_objective(out, mode, [w], calls[i]["output_steps"],
                                       monitor_x=mon_x)

  def leaves():
    return [e.detach().clone().requires_grad_(True) for e in base_eps]

  def batched():
    lv = leaves()
    outs = fdtdz_batch([dict(c, epsilon=lv[v]) for c, v in zip(calls, variant_of)])
    sum(objective(i, o) for i, o in enumerate(outs)).backward()
    return [e.grad for e in lv]

  def sequential(kernel):
    lv = leaves()
    for i, (c, v) in enumerate(zip(calls, variant_of)):
      out = fdtdz(**dict(c, epsilon=lv[v], launch_params=dict(kernel=kernel)))
      objective(i, out).backward()
    return [e.grad for e in lv]

  def timed(fn, *a):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
      r = fn(*a)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / args.steps, r

  for _ in range(args.warmup):
    batched()
    sequential("auto")
  tb, gb = timed(batched)
  ts, gs = timed(sequential, "auto")
  t1, _ = timed(sequential, "stream")
  tn, _ = timed(sequential, "reference")
  diff = max(float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
             for a, b in zip(gb, gs))
  shape = calls[0]["epsilon"].shape
  line = {"metric": METRIC, "value": round(cells * steps / tb / 1e9, 3),
          "unit": "Gcell-updates/s (forward steps per second of forward + gradient)",
          "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
          "ms_per_step": round(tb * 1e3, 3), "higher_is_better": True, "scaling": "weak",
          "vs_baseline": None, "dtype": "f32", "data": "synthetic",
          "config": {"workload": f"design_batch: 16 variants x 2 ports x 2 frequencies = "
                                 f"{len(calls)} simulations of {shape[1]}x{shape[2]}x{shape[3]} "
                                 f"cells, {steps} steps, forward + adjoint gradient of the "
                                 f"overlap objective (reduced from 384x384x128, 8000 steps)",
                     "cells": cells, "time_steps": steps, "batched_s": round(tb, 4),
                     "sequential_best_s": round(ts, 4),
                     "sequential_one_step_kernel_s": round(t1, 4),
                     "sequential_reference_kernel_s": round(tn, 4),
                     "speedup_vs_sequential_best": round(ts / tb, 3),
                     "speedup_vs_sequential_one_step_kernel": round(t1 / tb, 3),
                     "speedup_vs_sequential_reference_kernel": round(tn / tb, 3),
                     "grad_rel_l2_batched_vs_sequential": diff}}
  print(json.dumps(line), flush=True)


def bench_design_full(args):
  """The batched inverse-design iteration at its full size: 16 variants x 2
  ports x 2 frequencies = 64 simulations of 384x384x128 cells, 160x160x8
  design squares, 8000 steps, forward + adjoint gradient of the summed
  overlap objective.  The gradient workspaces (~60 GB a simulation at 8000
  steps) do not fit in HBM together, so the batch runs in waves
  (fdtdz_batch_value_and_grad); the baseline is 64 sequential calls of the
  default single-simulation path, each differentiated before the next."""
  import torch
  from pjz_amd import fdtdz, fdtdz_batch_value_and_grad, workloads
  from pjz_amd import fdtdz_torch as F
  steps = args.time_steps or 8000
  calls, variant_of, monitors = workloads.design_batch(steps=steps, xx=384, yy=384, zz=128,
                                                       design=160)
  nv = max(variant_of) + 1
  base_eps = [None] * nv
  for c, v in zip(calls, variant_of):
    base_eps[v] = c["epsilon"]
  cells = sum(int(c["epsilon"].shape[1] * c["epsilon"].shape[2] * c["epsilon"].shape[3])
              for c in calls)

  t_start = [time.perf_counter()]

  def objective(i, out):
    # Progress on stderr as each simulation's objective is formed (a wave's
    # forward is done); long runs would otherwise print nothing for minutes.
    print(f"  simulation {i} forward done at {time.perf_counter() - t_start[0]:.0f} s",
          file=sys.stderr, flush=True)
    w, mon_x, mode = monitors[i]
    return workloads.overlap_objective(out, mode, [w], calls[i]["output_steps"],
                                       monitor_x=mon_x)

  def leaves():
    return [e.detach().clone().requires_grad_(True) for e in base_eps]

  waves = []

  def batched():
    lv = leaves()
    run = [dict(c, epsilon=lv[v]) for c, v in zip(calls, variant_of)]
    probs = F._batch_problems(run)[0]
    S = F.batch_ckpt_every(probs)
    for p in probs:
      p.ckpt_every = S
    free, _ = torch.cuda.mem_get_info()
    budget = int(0.85 * (free + torch.cuda.memory_reserved() - torch.cuda.memory_allocated()))
    waves[:] = F.plan_waves(probs, budget)
    fdtdz_batch_value_and_grad(run, objective, memory_budget=budget)
    return [e.grad for e in lv]

  def sequential():
    lv = leaves()
    for i, (c, v) in enumerate(zip(calls, variant_of)):
      out = fdtdz(**dict(c, epsilon=lv[v]))
      objective(i, out).backward()
      del out
    return [e.grad for e in lv]

  def timed(fn):
    torch.cuda.synchronize()
    t0 = t_start[0] = time.perf_counter()
    for _ in range(args.steps):
      r = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / args.steps, r

  for _ in range(args.warmup):
    batched()
  tb, gb = timed(batched)
  print(f"design_batch_384: batched {tb:.1f} s in {len(waves)} waves", file=sys.stderr,
        flush=True)
  ts, gs = timed(sequential)
  diff = max(float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
             for a, b in zip(gb, gs))
  line = {"metric": METRIC, "value": round(cells * steps / tb / 1e9, 3),
          "unit": "Gcell-updates/s (forward steps per second of forward + gradient)",
          "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
          "ms_per_step": round(tb * 1e3, 3), "higher_is_better": True, "scaling": "weak",
          "vs_baseline": None, "dtype": "f32", "data": "synthetic",
          "config": {"workload": f"design_batch_384: 16 variants x 2 ports x 2 frequencies = "
                                 f"{len(calls)} simulations of 384x384x128 cells, 160x160x8 "
                                 f"design squares, {steps} steps, forward + adjoint gradient "
                                 f"of the overlap objective",
                     "cells": cells, "time_steps": steps, "batched_s": round(tb, 3),
                     "waves": [len(w) for w in waves],
                     "sequential_best_s": round(ts, 3),
                     "speedup_vs_sequential_best": round(ts / tb, 3),
                     "grad_rel_l2_batched_vs_sequential": diff}}
  print(json.dumps(line), flush=True)


def bench_mode(args):
  """pjz's mode solver: microseconds per subspace iteration of the HIP kernel
  (one launch runs the loop) against the same iteration as torch ops on the
  same GPU, on the 30 x 20 strip of pjz's tests (4 modes) and a 256 x 128
  section (2 modes), `tol = 0` so both run exactly the iteration count."""
  import numpy as np
  import torch
  from pjz_amd import _mode as M
  dev = torch.device("cuda", 0)
  iters = args.time_steps or 2000
  rows = []
  for (uu, vv, mm) in ((30, 20, 4), (256, 128, 2)):
    eps = np.ones((3, uu, vv))
    eps[:, uu * 3 // 10: uu * 7 // 10, vv * 2 // 5: vv * 3 // 5] = 12.25
    e = torch.as_tensor(eps[:, None], device=dev)
    omega = np.array([2 * np.pi / 37])
    res = {}
    for impl in ("hip", "torch"):
      limit = M.HIP_MAX_MODES
      if impl == "torch":
        M.HIP_MAX_MODES = 0
      try:
        M.mode(e, omega, num_modes=mm, max_iters=20)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        beta, _, _, it = M.mode(e, omega, num_modes=mm, max_iters=iters, tol=0.0)
        torch.cuda.synchronize()
        res[impl] = ((time.perf_counter() - t0) / it * 1e6, beta[0].cpu().numpy())
      finally:
        M.HIP_MAX_MODES = limit
    rows.append({"section": f"{uu}x{vv}", "modes": mm, "iterations": iters,
                 "hip_us_per_iter": round(res["hip"][0], 2),
                 "torch_us_per_iter": round(res["torch"][0], 2),
                 "beta_max_rel_diff": float(np.max(np.abs(res["hip"][1] - res["torch"][1]) /
                                                   np.abs(res["torch"][1])))})
  line = {"metric": "mode solver subspace iterations per second", "unit": "iterations/s",
          "value": round(1e6 / rows[0]["hip_us_per_iter"], 1), "n_gpus": 1, "steps": 1,
          "warmup": 1, "ms_per_step": round(rows[0]["hip_us_per_iter"] * iters / 1e3, 3),
          "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
          "data": "synthetic", "config": {"workload": "mode: pjz mode() subspace iteration",
                                          "cases": rows}}
  print(json.dumps(line), flush=True)


def bench_density(args):
  """pjz's density() on the HIP kernels: forward + reverse pass of a
  4096 x 4096 design variable (cone radius 4: 81 taps), against the torch
  restatement of the same function on the same GPU (float32), with the blur
  kernel's arithmetic rate beside the FP32 vector peak."""
  import numpy as np
  import torch
  from pjz_amd import _density as D
  dev = torch.device("cuda", 0)
  X = Y = 4096
  radius, alpha = 4.0, 0.6
  reps = args.steps
  u = torch.rand((X, Y), device=dev)
  gd, gl = torch.randn((X, Y), device=dev), torch.randn((X, Y), device=dev)

  def hip():
    v = u.clone().requires_grad_(True)
    dens, loss = D.density(v, radius, alpha)
    ((dens * gd).sum() + (loss * gl).sum()).backward()

  def restated():
    v = u.clone().requires_grad_(True)
    uf = D.blur(v, radius)
    up = D.project(uf, 0.5)
    dens, loss = alpha * up + (1 - alpha) * uf, D.geom_loss(up, uf, 1.0, 0.25, 0.75)
    ((dens * gd).sum() + (loss * gl).sum()).backward()

  times = {}
  for name, fn in (("hip", hip), ("torch", restated)):
    for _ in range(max(args.warmup, 1)):
      fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
      fn()
    torch.cuda.synchronize()
    times[name] = (time.perf_counter() - t0) / reps
  # The forward alone (blur + projection kernels), events on the current
  # stream; the blur's multiply-adds over that time bound its rate below.
  v = u.clone()
  taps = D.cone(radius).shape[0] ** 2
  start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
  start.record()
  for _ in range(reps):
    D.density(v, radius, alpha)
  end.record()
  torch.cuda.synchronize()
  fwd_s = start.elapsed_time(end) / 1e3 / reps
  line = {"metric": "density() forward + gradient, pixels per second", "unit": "Mpixel/s",
          "value": round(X * Y / times["hip"] / 1e6, 1), "n_gpus": 1, "steps": reps,
          "warmup": args.warmup, "ms_per_step": round(times["hip"] * 1e3, 3),
          "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
          "data": "synthetic",
          "config": {"workload": f"density: {X}x{Y} variable, cone radius {radius} "
                                 f"({taps} taps), forward + reverse pass",
                     "hip_ms": round(times["hip"] * 1e3, 3),
                     "torch_restated_ms": round(times["torch"] * 1e3, 3),
                     "speedup_vs_torch": round(times["torch"] / times["hip"], 2),
                     "forward_ms": round(fwd_s * 1e3, 3),
                     "forward_blur_tflops_bound": round(2 * taps * X * Y / fwd_s / 1e12, 2),
                     "fp32_vector_peak_tflops": 157.3}}
  print(json.dumps(line), flush=True)


def bench_render(args):
  """pjz's epsilon() on the HIP render kernels: forward + reverse pass of a
  3-layer stack magnified 2m = 4 times onto 384 x 384 x 128 cells (the
  inverse-design size), against the torch restatement of the same function
  on the same GPU (float32)."""
  import torch
  from pjz_amd import _epsilon as E
  dev = torch.device("cuda", 0)
  xx = yy = 384
  zz, m = 128, 2
  gen = torch.Generator(device=dev).manual_seed(0)
  layers = 1 + 11 * torch.rand((3, 2 * m * xx, 2 * m * yy), generator=gen, device=dev)
  pos = torch.tensor([40.3, 55.7], device=dev)
  g = torch.randn((3, xx, yy, zz), generator=gen, device=dev)
  k = torch.arange(zz, dtype=torch.float32, device=dev)[:, None]
  gs = k + torch.tensor([[-0.5, 0.0]], device=dev)
  ge = k + torch.tensor([[0.5, 1.0]], device=dev)

  def hip():
    lv, pv = layers.clone().requires_grad_(True), pos.clone().requires_grad_(True)
    (E.render(lv, pv, gs, ge, m) * g).sum().backward()

  def restated():
    lv, pv = layers.clone().requires_grad_(True), pos.clone().requires_grad_(True)
    out = torch.stack([E._render_single(lv, pv, gs, ge, m, a, False) for a in "xyz"])
    (out * g).sum().backward()

  times = {}
  for name, fn in (("hip", hip), ("torch", restated)):
    for _ in range(max(args.warmup, 1)):
      fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
      fn()
    torch.cuda.synchronize()
    times[name] = (time.perf_counter() - t0) / args.steps
  cells = 3 * xx * yy * zz
  line = {"metric": "epsilon() render forward + gradient, Yee-cell components per second",
          "unit": "Mcomponent/s", "value": round(cells / times["hip"] / 1e6, 1), "n_gpus": 1,
          "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(times["hip"] * 1e3, 3),
          "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
          "data": "synthetic",
          "config": {"workload": f"render: 3 layers of {2 * m * xx}x{2 * m * yy} sub-pixels "
                                 f"(2m = {2 * m}) -> (3, {xx}, {yy}, {zz}), forward + reverse",
                     "hip_ms": round(times["hip"] * 1e3, 3),
                     "torch_restated_ms": round(times["torch"] * 1e3, 3),
                     "speedup_vs_torch": round(times["torch"] / times["hip"], 2)}}
  print(json.dumps(line), flush=True)


def main():
  ap = argparse.ArgumentParser()
  ap.add_argument("--gpus", type=int, default=1)
  ap.add_argument("--steps", type=int, default=3)
  ap.add_argument("--warmup", type=int, default=1)
  ap.add_argument("--workload", default="native_512")
  ap.add_argument("--kernel", default="auto",
                  choices=["auto", "reference", "stream", "stream2", "stream3", "stream4"])
  ap.add_argument("--time-steps", type=int, default=0,
                  help="override the FDTD step count (profiling runs only)")
  ap.add_argument("--no-cpu-baseline", action="store_true")
  args = ap.parse_args()
  if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
    print("bench.py measures one GPU: the simulation is a single-GPU kernel and "
          "multi-GPU runs are not part of this build (use --gpus 1)", file=sys.stderr)
    sys.exit(2)
  if args.workload == "grad_256":
    return bench_grad(args)
  if args.workload == "batch64":
    return bench_batch(args)
  if args.workload == "design_batch":
    return bench_design(args)
  if args.workload == "design_batch_384":
    return bench_design_full(args)
  if args.workload == "mode":
    return bench_mode(args)
  if args.workload == "density":
    return bench_density(args)
  if args.workload == "render":
    return bench_render(args)

  import numpy as np
  import torch
  from pjz_amd import fdtdz, workloads

  builder = getattr(workloads, args.workload)
  kw = builder(steps=args.time_steps) if args.time_steps else builder()
  dev = torch.device("cuda", 0)
  for k, v in list(kw.items()):
    if isinstance(v, np.ndarray):
      kw[k] = torch.as_tensor(v).to(dev)
  kw["launch_params"] = dict(kernel=args.kernel)
  cells = workloads.cells(kw)
  tsteps = workloads.time_steps(kw)
  lo, hi = kw["pml_widths"]
  zz = kw["pml_sigma"].shape[0]

  for _ in range(args.warmup):
    fdtdz(**kw)
  torch.cuda.synchronize()
  stream = torch.cuda.current_stream(dev)
  starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
  ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
  t0 = time.perf_counter()
  for i in range(args.steps):
    starts[i].record(stream)
    fdtdz(**kw)
    ends[i].record(stream)
  torch.cuda.synchronize()
  wall = time.perf_counter() - t0
  gpu_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
  ms_per_step = wall * 1e3 / args.steps
  value = cells * tsteps * args.steps / wall / 1e9

  kernel_name = resolved_stepper(kw)
  # Stepper launches per run: a kernel of depth T advances T steps per
  # launch; the remainder runs as one two-step and/or one one-step launch.
  # The reference stepper's H + E launch pair counts as one launch of one
  # step.
  temporal_block = {"stream2": 2, "stream3": 3, "stream4": 4}.get(kernel_name, 1)
  rem = tsteps % temporal_block
  launches = tsteps // temporal_block + (rem // 2 + rem % 2)
  bpc = algorithmic_bytes_per_cell(zz, lo, hi)
  bytes_per_launch = cells * bpc
  avg_launch_s = (sum(gpu_ms) / len(gpu_ms)) / 1e3 / launches
  achieved = bytes_per_launch / avg_launch_s / 1e9

  line = {
      "metric": METRIC, "value": round(value, 3), "unit": "Gcell-updates/s",
      "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
      "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak",
      "vs_baseline": None, "dtype": "f32", "data": "synthetic",
      "config": {"workload": f"{args.workload}: {kw['epsilon'].shape[1]}x"
                             f"{kw['epsilon'].shape[2]}x{kw['epsilon'].shape[3]} cells, "
                             f"{tsteps} FDTD steps, "
                             f"{len(range(*kw['output_steps']))} snapshots",
                 "cells": int(cells), "time_steps": int(tsteps), "kernel": kernel_name,
                 "temporal_block": temporal_block, "pml_widths": [int(lo), int(hi)]},
      "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                   "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                   "traffic": measured_traffic(kernel_name, args.workload),
                   "bytes_per_launch": round(bytes_per_launch),
                   "algorithmic_bytes_per_cell": round(bpc, 4),
                   "avg_launch_us": round(avg_launch_s * 1e6, 2)},
  }
  # The practical ceiling: a nontemporal 1 GiB device copy on the same GPU,
  # and the kernel's measured HBM traffic rate as a fraction of it.
  ceiling = copy_ceiling_gbs(dev)
  line["roofline"]["copy_ceiling_gbs"] = round(ceiling, 1)
  traffic = line["roofline"]["traffic"]
  if traffic:
    line["roofline"]["traffic_frac_of_copy"] = round(traffic / avg_launch_s / 1e9 / ceiling, 4)
  if args.kernel == "auto" and not args.time_steps:
    line["one_step_kernel"] = one_step_roofline(builder, dev, bytes_per_launch, cells)
  if not args.no_cpu_baseline:
    np_kw = builder(steps=args.time_steps) if args.time_steps else builder()
    line["cpu_baseline"] = cpu_baseline(np_kw)
  print(json.dumps(line), flush=True)


def resolved_stepper(kw):
  """Name of the stepper kernel fdtdz() runs for these arguments on this
  device (the library resolves "auto" / "autotune")."""
  from pjz_amd import _lib
  from pjz_amd import fdtdz_torch as F
  shape = {k: tuple(v.shape) for k, v in kw.items() if hasattr(v, "shape")}
  prob = F.make_problem(shape["epsilon"], kw["dt"], shape["source_field"],
                        kw["source_position"], shape["absorption_mask"], shape["pml_kappa"],
                        kw["pml_widths"], kw["output_steps"], shape["source_waveform"][0],
                        kw["use_reduced_precision"], kw["launch_params"], kw["offset"])
  names = {_lib.KERNEL_REFERENCE: "reference", _lib.KERNEL_STREAM: "stream",
           _lib.KERNEL_STREAM2: "stream2", _lib.KERNEL_STREAM3: "stream3",
           _lib.KERNEL_STREAM4: "stream4"}
  return names[_lib.stepper(prob)]


def one_step_roofline(builder, dev, bytes_per_launch, cells, tsteps=1000):
  """The one-step kernel (k_stream, one HBM pass per FDTD step) on the same
  domain for `tsteps` steps: its roofline fraction beside the two-step
  kernel's, which does more cell-updates per second at a lower fraction."""
  import numpy as np
  import torch
  from pjz_amd import fdtdz, workloads

  kw = builder(steps=tsteps)
  for k, v in list(kw.items()):
    if isinstance(v, np.ndarray):
      kw[k] = torch.as_tensor(v).to(dev)
  kw["launch_params"] = dict(kernel="stream")
  n = workloads.time_steps(kw)
  stream = torch.cuda.current_stream(dev)
  fdtdz(**kw)
  torch.cuda.synchronize()
  start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
  start.record(stream)
  fdtdz(**kw)
  end.record(stream)
  torch.cuda.synchronize()
  launch_s = start.elapsed_time(end) / 1e3 / n
  achieved = bytes_per_launch / launch_s / 1e9
  return {"kernel": "stream", "time_steps": int(n), "avg_launch_us": round(launch_s * 1e6, 2),
          "gcell_updates_per_s": round(cells / launch_s / 1e9, 3),
          "achieved": round(achieved, 1), "frac": round(achieved / HBM_PEAK_GBS, 4)}


if __name__ == "__main__":
  main()
