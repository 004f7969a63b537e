"""``fdtdz()``: the drop-in for the kernel call pjz makes.

pjz calls ``fdtdz_jax.fdtdz(**kwargs)`` at ``src/pjz/_field.py:254-269`` with
exactly the keywords of :func:`fdtdz` below.  This module keeps that keyword
set, the array shapes pjz passes and the ``(nn, 3, xx, yy, zz)`` return shape
(``_field.py:189-191,276-279``), and runs the simulation through
``libpjzfdtd.so`` on the current HIP stream.  Arrays may be torch tensors or
numpy arrays; the result is a float32 torch tensor on the GPU.  When
``epsilon`` requires grad, the call is a ``torch.autograd.Function`` whose
backward runs the discrete adjoint kernels (``pjz_fdtd_backward``).
"""

import ctypes
import math
from collections.abc import Mapping

import numpy as np
import torch

from . import _lib

_KERNELS = {"auto": _lib.KERNEL_AUTO, "reference": _lib.KERNEL_REFERENCE,
            "stream": _lib.KERNEL_STREAM, "stream2": _lib.KERNEL_STREAM2,
            "stream3": _lib.KERNEL_STREAM3, "stream4": _lib.KERNEL_STREAM4,
            "autotune": _lib.KERNEL_AUTOTUNE}
_LAUNCH_KEYS = {"kernel", "tile_x", "tile_y", "waves_per_simd", "ckpt_every", "ckpt_inner",
                "memory_budget", "graph_steps"}


def _launch_options(launch_params):
  """Parses ``launch_params``: None, a mapping, or a tuple of (key, value).

  Keys: ``kernel`` ("auto" | "reference" | "stream" | "stream2" | "stream3" |
  "stream4" (one, two, three or four steps per pass) | "autotune": the first
  call per shape times the kernels, depths and tilings, later calls reuse the
  fastest), ``tile_x``, ``tile_y``,
  ``waves_per_simd`` (stream kernel tiling), ``ckpt_every`` (steps between
  checkpoints kept for the gradient; default about sqrt(2 * steps)),
  ``ckpt_inner`` (single calls: steps between the second level of
  checkpoints the backward keeps inside each ``ckpt_every`` segment; 0 = one
  level), ``memory_budget`` (single calls with a gradient: bytes the
  workspace may take; default 90% of the device's free memory.  Without an
  explicit ``ckpt_every``, checkpoints go to two levels when one does not
  fit), ``graph_steps`` (batches: steps per captured and replayed hipGraph
  chunk, even, <= 128; 0, the default, launches every step pair directly).
  """
  if launch_params is None:
    return {}
  if isinstance(launch_params, Mapping):
    opts = dict(launch_params)
  else:
    try:
      opts = dict(launch_params)
    except (TypeError, ValueError):
      raise ValueError(f"launch_params must be None, a mapping or (key, value) pairs, "
                       f"got {launch_params!r}") from None
  unknown = set(opts) - _LAUNCH_KEYS
  if unknown:
    raise ValueError(f"unknown launch_params keys {sorted(unknown)}")
  if "kernel" in opts and opts["kernel"] not in _KERNELS:
    raise ValueError(f"launch_params kernel must be one of {sorted(_KERNELS)}")
  return opts


def source_geometry(source_shape, domain):
  """(axis, quadratures) of a source plane from its shape; ValueError if bad.

  x plane (2, 1, YY, ZZ), y plane (2, XX, 1, ZZ), z plane (2, 2, XX, YY, 1)
  with two quadratures or (2, XX, YY, 1) with one (``_field.py:206-236``).
  """
  XX, YY, ZZ = domain
  shp = tuple(int(s) for s in source_shape)
  if len(shp) == 5:
    if shp[0] not in (1, 2) or shp[1:] != (2, XX, YY, 1):
      raise ValueError(f"z-plane source_field must be (2, 2, {XX}, {YY}, 1), got {shp}")
    return 2, shp[0]
  if len(shp) != 4 or shp[0] != 2:
    raise ValueError(f"source_field must have shape (2, ...) with 4 dims or "
                     f"(2, 2, ...) with 5 dims, got {shp}")
  expect = {0: (2, 1, YY, ZZ), 1: (2, XX, 1, ZZ), 2: (2, XX, YY, 1)}
  matches = [a for a in (0, 1, 2) if shp == expect[a]]
  if not matches:
    raise ValueError(f"source_field shape {shp} is none of {list(expect.values())}")
  if len(matches) > 1:
    raise ValueError(f"source_field shape {shp} is ambiguous for domain {domain}")
  return matches[0], 1


def make_problem(epsilon_shape, dt, source_shape, source_position, mask_shape,
                 pml_shape, pml_widths, output_steps, tt, use_reduced_precision,
                 launch_params, offset):
  """Builds the C ``pjz_fdtd_problem`` from static arguments and shapes."""
  eps_shape = tuple(int(s) for s in epsilon_shape)
  if len(eps_shape) != 4 or eps_shape[0] != 3:
    raise ValueError(f"epsilon must be (3, xx, yy, zz), got {eps_shape}")
  mask_shape = tuple(int(s) for s in mask_shape)
  if len(mask_shape) != 3 or mask_shape[0] != 3:
    raise ValueError(f"absorption_mask must be (3, XX, YY), got {mask_shape}")
  pml_shape = tuple(int(s) for s in pml_shape)
  if len(pml_shape) != 2 or pml_shape[1] != 2:
    raise ValueError(f"pml arrays must be (ZZ, 2), got {pml_shape}")
  domain = (mask_shape[1], mask_shape[2], pml_shape[0])
  axis, quad = source_geometry(source_shape, domain)
  if len(tuple(output_steps)) != 3:
    raise ValueError("output_steps must be (start, stop, step)")
  if len(tuple(pml_widths)) != 2 or len(tuple(offset)) != 3:
    raise ValueError("pml_widths must be (lo, hi) and offset (x0, y0, z0)")
  opts = _launch_options(launch_params)
  p = _lib.Problem()
  p.xx, p.yy, p.zz = domain
  p.ex, p.ey, p.ez = eps_shape[1:]
  p.x0, p.y0, p.z0 = (int(v) for v in offset)
  p.pml_lo, p.pml_hi = (int(v) for v in pml_widths)
  p.src_axis, p.src_pos, p.src_quad = axis, int(source_position), quad
  p.tt = int(tt)
  p.out_start, p.out_stop, p.out_step = (int(v) for v in output_steps)
  p.dt = float(dt)
  p.flags = 1 if use_reduced_precision else 0
  p.kernel = _KERNELS[opts.get("kernel", "auto")]
  p.tile_x = int(opts.get("tile_x", 0))
  p.tile_y = int(opts.get("tile_y", 0))
  p.waves_per_simd = int(opts.get("waves_per_simd", 0))
  p.ckpt_every = int(opts.get("ckpt_every", 0))
  p.ckpt_inner = int(opts.get("ckpt_inner", 0))
  p.graph_steps = int(opts.get("graph_steps", 0))
  return p


def steps_of(prob):
  """FDTD steps a call runs: the last snapshot step + 1 (0 if none)."""
  r = range(prob.out_start, prob.out_stop, prob.out_step)
  return r[-1] + 1 if len(r) else 0


def default_ckpt_every(nsteps):
  """Checkpoint spacing that balances checkpoints (~24 B/cell each) against
  the per-step E of one segment (12 B/cell each): about sqrt(2 * steps)."""
  return max(1, min(nsteps, int(round(math.sqrt(2.0 * max(nsteps, 1))))))


def _with_ckpt(prob, every, inner):
  q = _lib.Problem()
  ctypes.pointer(q)[0] = prob
  q.ckpt_every, q.ckpt_inner = every, inner
  return q


def plan_checkpoints(prob, budget):
  """(ckpt_every, ckpt_inner) for a single call's gradient within ``budget``
  workspace bytes: one level every ~sqrt(2 steps) steps if that fits, else
  the two-level spacing with the smallest workspace (one more forward pass in
  the backward).  Raises ValueError if no spacing fits."""
  T = steps_of(prob)
  one = default_ckpt_every(T)
  if _lib.workspace_bytes(_with_ckpt(prob, one, 0)) <= budget:
    return one, 0
  # Workspace = fixed part + (ceil(T/S1) + ceil(S1/S2)) states + (S2 + 1)
  # E changes: search the inner spacing near its optimum, the outer one
  # near sqrt(T S2) for each.
  best = None
  for s2 in range(1, max(2, int(3 * T ** (1 / 3.0)) + 2)):
    c = int(round(math.sqrt(T * s2)))
    for s1 in {max(s2 + 1, c + d) for d in range(-2 * s2, 2 * s2 + 1, max(1, s2 // 2))}:
      if s1 >= T:
        continue
      w = _lib.workspace_bytes(_with_ckpt(prob, s1, s2))
      if best is None or w < best[0]:
        best = (w, s1, s2)
  if best is None or best[0] > budget:
    need = best[0] if best else _lib.workspace_bytes(_with_ckpt(prob, one, 0))
    raise ValueError(f"the gradient workspace needs at least {need} bytes with two levels of "
                     f"checkpoints, above the memory budget of {int(budget)} bytes")
  return best[1], best[2]


def _device_of(arrays):
  for a in arrays:
    if isinstance(a, torch.Tensor) and a.is_cuda:
      return a.device
  if not torch.cuda.is_available():
    raise RuntimeError("fdtdz needs a HIP GPU: torch.cuda.is_available() is False")
  return torch.device("cuda", torch.cuda.current_device())


def _as_dev(a, device):
  if isinstance(a, torch.Tensor):
    t = a.detach()
  else:
    t = torch.as_tensor(np.asarray(a))
  return t.to(device=device, dtype=torch.float32).contiguous()


def _ptr(t):
  return ctypes.c_void_p(t.data_ptr())


class _Workspace:
  """Device scratch for one call, 256-byte aligned."""

  def __init__(self, nbytes, device):
    self.nbytes = nbytes
    self.buf = torch.empty(nbytes + 256, dtype=torch.uint8, device=device)
    self.ptr = ctypes.c_void_p((self.buf.data_ptr() + 255) & ~255)


def _forward(prob, dev, device):
  """Runs pjz_fdtd_forward; returns (out, workspace)."""
  nn = _lib.num_outputs(prob)
  out = torch.empty((nn, 3, prob.ex, prob.ey, prob.ez), dtype=torch.float32, device=device)
  ws = None
  if nn == 0:
    return out, ws
  ws = _Workspace(_lib.workspace_bytes(prob), device)
  with torch.cuda.device(device):
    stream = torch.cuda.current_stream(device).cuda_stream
    rc = _lib.load().pjz_fdtd_forward(
        ctypes.byref(prob), _ptr(dev[0]), _ptr(dev[3]), _ptr(dev[4]), _ptr(dev[5]),
        _ptr(dev[6]), _ptr(dev[1]), _ptr(dev[2]), _ptr(out), ws.ptr,
        ctypes.c_size_t(ws.nbytes), ctypes.c_void_p(stream))
  _lib.check(rc, "pjz_fdtd_forward")
  # dev and ws were allocated on this stream, so the caching allocator only
  # reuses their memory for work ordered after the kernels just enqueued.
  return out, ws


def _backward(prob, dev, device, ws, grad_out):
  grad_eps = torch.empty((3, prob.ex, prob.ey, prob.ez), dtype=torch.float32, device=device)
  if _lib.num_outputs(prob) == 0:  # no steps run: nothing depends on epsilon
    return grad_eps.zero_()
  g = grad_out.detach().to(device=device, dtype=torch.float32).contiguous()
  with torch.cuda.device(device):
    stream = torch.cuda.current_stream(device).cuda_stream
    rc = _lib.load().pjz_fdtd_backward(
        ctypes.byref(prob), _ptr(dev[0]), _ptr(dev[3]), _ptr(dev[4]), _ptr(dev[5]),
        _ptr(dev[6]), _ptr(dev[1]), _ptr(dev[2]), _ptr(g), _ptr(grad_eps), ws.ptr,
        ctypes.c_size_t(ws.nbytes), ctypes.c_void_p(stream))
  _lib.check(rc, "pjz_fdtd_backward")
  return grad_eps


class FdtdzFunction(torch.autograd.Function):
  """fdtdz() with a gradient for ``epsilon`` from the adjoint kernels.

  The checkpoint workspace is released after the first backward (it holds
  several full field states).  A later backward through the same graph
  (``retain_graph=True``) first reruns the forward simulation to rebuild it:
  the same gradient, for one more forward pass.
  """

  @staticmethod
  def forward(ctx, epsilon, prob, dev, device):
    out, ws = _forward(prob, dev, device)
    ctx.prob, ctx.dev, ctx.device, ctx.ws = prob, dev, device, ws
    ctx.eps_device, ctx.eps_dtype = epsilon.device, epsilon.dtype
    return out

  @staticmethod
  def backward(ctx, grad_out):
    ws = ctx.ws
    ctx.ws = None
    if ws is None and _lib.num_outputs(ctx.prob) > 0:
      ws = _forward(ctx.prob, ctx.dev, ctx.device)[1]
    grad = _backward(ctx.prob, ctx.dev, ctx.device, ws, grad_out)
    return grad.to(device=ctx.eps_device, dtype=ctx.eps_dtype), None, None, None


def fdtdz(epsilon, dt, source_field, source_waveform, source_position,
          absorption_mask, pml_kappa, pml_sigma, pml_alpha, pml_widths,
          output_steps, use_reduced_precision, launch_params, offset=(0, 0, 0)):
  """Runs the FDTD simulation; returns E snapshots (nn, 3, xx, yy, zz) float32.

  Same keywords, shapes and meaning as the call at ``src/pjz/_field.py:254-269``.
  Snapshot i is E after time step ``output_steps[0] + i * output_steps[2]``
  over the epsilon window at ``offset``.  Differentiable with respect to
  ``epsilon`` when it is a tensor that requires grad.  Raises ValueError for
  inconsistent shapes or arguments, RuntimeError for GPU failures.
  """
  arrays = (epsilon, source_field, source_waveform, absorption_mask,
            pml_kappa, pml_sigma, pml_alpha)
  shapes = [tuple(np.shape(a)) if not isinstance(a, torch.Tensor) else tuple(a.shape)
            for a in arrays]
  for name, shp in zip(("pml_sigma", "pml_alpha"), shapes[5:]):
    if shp != shapes[4]:
      raise ValueError(f"{name} shape {shp} differs from pml_kappa {shapes[4]}")
  wshape = shapes[2]
  if len(wshape) != 2 or wshape[1] != 2:
    raise ValueError(f"source_waveform must be (tt, 2), got {wshape}")
  prob = make_problem(shapes[0], dt, shapes[1], source_position, shapes[3], shapes[4],
                      pml_widths, output_steps, wshape[0], use_reduced_precision,
                      launch_params, offset)
  _lib.num_outputs(prob)  # validates before any device work
  device = _device_of(arrays)
  _lib.load()
  dev = [_as_dev(a, device) for a in arrays]
  wants_grad = (torch.is_grad_enabled() and isinstance(epsilon, torch.Tensor)
                and epsilon.requires_grad)
  if not wants_grad:
    prob.ckpt_every = prob.ckpt_inner = 0
    return _forward(prob, dev, device)[0]
  if prob.ckpt_every <= 0:
    opts = _launch_options(launch_params)
    budget = opts.get("memory_budget")
    if budget is None:
      free, _ = torch.cuda.mem_get_info(device)
      cached = torch.cuda.memory_reserved(device) - torch.cuda.memory_allocated(device)
      budget = 0.9 * (free + cached)
    prob.ckpt_every, prob.ckpt_inner = plan_checkpoints(prob, float(budget))
  return FdtdzFunction.apply(epsilon, prob, dev, device)


def _batch_problems(calls):
  """(probs ctypes array, per-call device arrays, device) of a batch."""
  n = len(calls)
  probs = (_lib.Problem * n)()
  devs = []
  device = None
  for i, kw in enumerate(calls):
    kw = dict(kw)
    kw.setdefault("offset", (0, 0, 0))
    arrays = (kw["epsilon"], kw["source_field"], kw["source_waveform"], kw["absorption_mask"],
              kw["pml_kappa"], kw["pml_sigma"], kw["pml_alpha"])
    shapes = [tuple(a.shape) if isinstance(a, torch.Tensor) else tuple(np.shape(a))
              for a in arrays]
    if len(shapes[2]) != 2 or shapes[2][1] != 2:
      raise ValueError(f"call {i}: source_waveform must be (tt, 2), got {shapes[2]}")
    probs[i] = make_problem(shapes[0], kw["dt"], shapes[1], kw["source_position"], shapes[3],
                            shapes[4], kw["pml_widths"], kw["output_steps"], shapes[2][0],
                            kw.get("use_reduced_precision", False), kw.get("launch_params"),
                            kw["offset"])
    _lib.num_outputs(probs[i])
    device = device or _device_of(arrays)
    devs.append([_as_dev(a, device) for a in arrays])
  return probs, devs, device


def _batch_workspace_bytes(probs):
  lib = _lib.load()
  nbytes = ctypes.c_size_t(0)
  _lib.check(lib.pjz_fdtd_batch_workspace_bytes(probs, len(probs), ctypes.byref(nbytes)),
             "pjz_fdtd_batch_workspace_bytes")
  return nbytes.value


def _batch_workspace(probs, device):
  return _Workspace(_batch_workspace_bytes(probs), device)


def _subset(probs, idx, ckpt_every=None):
  """ctypes array of the problems ``idx`` (copies; ``ckpt_every`` if given)."""
  sub = (_lib.Problem * len(idx))()
  for j, i in enumerate(idx):
    ctypes.memmove(ctypes.byref(sub[j]), ctypes.byref(probs[i]), ctypes.sizeof(_lib.Problem))
    if ckpt_every is not None:
      sub[j].ckpt_every = ckpt_every
      if ckpt_every == 0:
        sub[j].ckpt_inner = 0
  return sub


def plan_waves(probs, budget):
  """Splits a differentiated batch into consecutive waves whose gradient
  workspaces (checkpoints, per-step E changes of one segment, adjoint state)
  each fit in ``budget`` bytes; a list of index lists."""
  waves, cur = [], []
  for i in range(len(probs)):
    if cur and _batch_workspace_bytes(_subset(probs, cur + [i])) > budget:
      waves.append(cur)
      cur = []
    if not cur and _batch_workspace_bytes(_subset(probs, [i])) > budget:
      raise ValueError(f"batch item {i}: its gradient workspace of "
                       f"{_batch_workspace_bytes(_subset(probs, [i])) / 2**30:.1f} GiB exceeds the "
                       f"memory budget of {budget / 2**30:.1f} GiB")
    cur.append(i)
  if cur:
    waves.append(cur)
  return waves


def default_memory_budget(probs, device):
  """Device memory a differentiated batch may give its workspaces: 90% of
  what is free (including torch's cached blocks) less room for the outputs
  and for their cotangents, which exist alongside the workspace during the
  backward."""
  free, _ = torch.cuda.mem_get_info(device)
  cached = torch.cuda.memory_reserved(device) - torch.cuda.memory_allocated(device)
  outs = sum(4 * _lib.num_outputs(p) * 3 * p.ex * p.ey * p.ez for p in probs)
  return int(0.9 * (free + cached)) - 2 * outs


def _ptrs(tensors):
  return (ctypes.c_void_p * len(tensors))(*[t.data_ptr() for t in tensors])


def _batch_forward(probs, devs, device):
  """Runs pjz_fdtd_batch_forward; returns (outs, workspace)."""
  n = len(probs)
  outs = [torch.empty((_lib.num_outputs(p), 3, p.ex, p.ey, p.ez), dtype=torch.float32,
                      device=device) for p in probs]
  ws = _batch_workspace(probs, device)
  col = lambda j: _ptrs([d[j] for d in devs])
  with torch.cuda.device(device):
    stream = torch.cuda.current_stream(device).cuda_stream
    rc = _lib.load().pjz_fdtd_batch_forward(
        probs, n, col(0), col(3), col(4), col(5), col(6), col(1), col(2), _ptrs(outs), ws.ptr,
        ctypes.c_size_t(ws.nbytes), ctypes.c_void_p(stream))
  _lib.check(rc, "pjz_fdtd_batch_forward")
  return outs, ws


def _batch_backward(probs, devs, device, ws, grad_outs):
  n = len(probs)
  grads = [torch.empty((3, p.ex, p.ey, p.ez), dtype=torch.float32, device=device) for p in probs]
  gouts = [torch.zeros((_lib.num_outputs(p), 3, p.ex, p.ey, p.ez), dtype=torch.float32,
                       device=device) if g is None else
           g.detach().to(device=device, dtype=torch.float32).contiguous()
           for p, g in zip(probs, grad_outs)]
  col = lambda j: _ptrs([d[j] for d in devs])
  with torch.cuda.device(device):
    stream = torch.cuda.current_stream(device).cuda_stream
    rc = _lib.load().pjz_fdtd_batch_backward(
        probs, n, col(0), col(3), col(4), col(5), col(6), col(1), col(2), _ptrs(gouts),
        _ptrs(grads), ws.ptr, ctypes.c_size_t(ws.nbytes), ctypes.c_void_p(stream))
  _lib.check(rc, "pjz_fdtd_batch_backward")
  return grads


class FdtdzBatchFunction(torch.autograd.Function):
  """A batch of fdtdz() calls with gradients for every ``epsilon``.

  ``waves`` splits the batch into groups whose gradient workspaces fit in
  device memory one at a time.  The forward advances every wave; only the
  last keeps its checkpoints (the others run without them).  The backward
  takes the waves last to first and reruns the forward of each wave but the
  last with checkpoints before its adjoint: one extra forward pass for all
  but the last wave, instead of every wave's checkpoints at once.  A later
  backward through the same graph (``retain_graph=True``) reruns every
  wave's forward.
  """

  @staticmethod
  def forward(ctx, probs, devs, device, waves, *epsilons):
    outs = [None] * len(probs)
    ws = None
    for k, wave in enumerate(waves):
      last = k == len(waves) - 1
      sub = _subset(probs, wave, ckpt_every=None if last else 0)
      wouts, wws = _batch_forward(sub, [devs[i] for i in wave], device)
      for i, o in zip(wave, wouts):
        outs[i] = o
      if last:
        ws = wws
      del wws
    ctx.probs, ctx.devs, ctx.device, ctx.waves, ctx.ws = probs, devs, device, waves, ws
    ctx.eps_meta = [(e.device, e.dtype) for e in epsilons]
    return tuple(outs)

  @staticmethod
  def backward(ctx, *grad_outs):
    ws = ctx.ws
    ctx.ws = None
    grads = [None] * len(ctx.probs)
    for k in range(len(ctx.waves) - 1, -1, -1):
      wave = ctx.waves[k]
      sub = _subset(ctx.probs, wave)
      devs = [ctx.devs[i] for i in wave]
      if ws is None:
        ws = _batch_forward(sub, devs, ctx.device)[1]
      wg = _batch_backward(sub, devs, ctx.device, ws, [grad_outs[i] for i in wave])
      ws = None
      for i, g in zip(wave, wg):
        grads[i] = g
    return (None, None, None, None) + tuple(
        g.to(device=d, dtype=t) if need else None
        for g, (d, t), need in zip(grads, ctx.eps_meta, ctx.needs_input_grad[4:]))


def batch_ckpt_every(probs):
  """Common even checkpoint spacing of a differentiated batch: the largest
  ``ckpt_every`` the calls ask for, else about sqrt(2 * steps) of the longest."""
  asked = max(p.ckpt_every for p in probs)
  s = asked if asked > 0 else default_ckpt_every(max(steps_of(p) for p in probs))
  return s + (s % 2)


def fdtdz_batch(calls, memory_budget=None):
  """Runs several fdtdz() calls, given as keyword dicts, as one batch.

  All simulations advance together through the batched kernels
  (``pjz_fdtd_batch_forward``): persistent two-step launches over a work queue
  of (simulation, tile) items.  Returns the list of snapshot tensors, in the
  order of ``calls``.  When any ``epsilon`` requires grad, the batch is one
  ``torch.autograd.Function`` whose backward runs the batched adjoint
  (``pjz_fdtd_batch_backward``) for every simulation at once.  Every call needs
  zz <= 64 or an even zz <= 128.  Tiles are chosen for the whole batch (x
  chunks and, at zz <= 64, 8 or 13 rows per tile, from a model of the work
  queue); ``launch_params`` ``tile_y`` = 8 or 13 on every zz <= 64 call fixes
  those rows, and ``graph_steps`` replays chunks of step pairs as hipGraphs.

  With gradients, ``memory_budget`` (bytes; default: 90% of the free device
  memory less room for the outputs and their cotangents) bounds the
  checkpoint workspace held at once: a batch whose workspace exceeds it runs
  in waves (:func:`plan_waves`), at the cost of one more forward pass for
  every wave but the last (:class:`FdtdzBatchFunction`).
  """
  if len(calls) == 0:
    return []
  probs, devs, device = _batch_problems(calls)
  epsilons = [kw["epsilon"] for kw in calls]
  wants_grad = torch.is_grad_enabled() and any(
      isinstance(e, torch.Tensor) and e.requires_grad for e in epsilons)
  if not wants_grad:
    for p in probs:
      p.ckpt_every = p.ckpt_inner = 0
    return _batch_forward(probs, devs, device)[0]
  S = batch_ckpt_every(probs)
  for p in probs:
    p.ckpt_every = S
  budget = memory_budget if memory_budget is not None else \
      default_memory_budget(probs, device)
  waves = plan_waves(probs, budget)
  eps_in = [e if isinstance(e, torch.Tensor) else torch.as_tensor(np.asarray(e))
            for e in epsilons]
  return list(FdtdzBatchFunction.apply(probs, devs, device, waves, *eps_in))


def fdtdz_batch_value_and_grad(calls, objective, memory_budget=None):
  """Value and epsilon-gradients of a separable objective over a batch.

  ``objective(i, out)`` returns the scalar term of call ``i`` from its
  snapshots.  The calls run in waves whose gradient workspaces fit in
  ``memory_budget`` bytes (default: 85% of the free device memory); each
  wave is one batched forward, its objective terms and one batched backward,
  which accumulates into the ``.grad`` of every epsilon that requires grad.
  Unlike one :func:`fdtdz_batch` over all calls, no wave's forward runs
  twice: a wave's snapshots are only needed by its own terms.  Returns the
  summed objective (a detached tensor).
  """
  if len(calls) == 0:
    return torch.zeros(())
  probs, _, device = _batch_problems(calls)
  S = batch_ckpt_every(probs)
  for p in probs:
    p.ckpt_every = S
  if memory_budget is None:
    free, _ = torch.cuda.mem_get_info(device)
    cached = torch.cuda.memory_reserved(device) - torch.cuda.memory_allocated(device)
    memory_budget = int(0.85 * (free + cached))
  total = None
  for wave in plan_waves(probs, memory_budget):
    sub = [dict(calls[i], launch_params=dict(calls[i].get("launch_params") or {},
                                             ckpt_every=S)) for i in wave]
    outs = fdtdz_batch(sub, memory_budget=1 << 62)
    loss = sum(objective(i, o) for i, o in zip(wave, outs))
    if loss.requires_grad:
      loss.backward()
    total = loss.detach() if total is None else total + loss.detach()
    del outs, loss
  return total
