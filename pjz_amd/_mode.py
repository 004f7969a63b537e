"""Waveguide modes on the GPU: pjz's ``mode()`` solver.

Mirrors ``src/pjz/_mode.py``: the waveguide operator of a 2-D cross-section
(``_operator``, ``_mode.py:22-50``, periodic forward/backward differences),
subspace iteration with a QR re-orthogonalisation every step
(``_subspace_iteration``, ``:99-142``), a shift equal to the largest
eigenvalue from ``shift_iters`` iterations on one vector, wavevector
``sqrt(w + shift)``, Poynting normalisation (``_poynting``, ``:72-76``) and
the axis / excitation conventions of ``mode()`` (``:145-257``).

On a CUDA ``epsilon`` with at most 10 modes the iteration is one launch of
the library's HIP kernel (``pjz_fdtd_mode_solve``: one workgroup per
frequency runs the whole loop, CholeskyQR2 for the QR); a missing library
raises.  CPU tensors, and more than 10 modes, run the same iteration as torch
ops where ``epsilon`` lives (QR by ``torch.linalg.qr``).  Results are float32
like pjz's.
"""

import ctypes

from typing import Optional

import numpy as np
import torch


def _diff(arr, axis, is_forward):
  if is_forward:
    return torch.roll(arr, -1, axis) - arr
  return arr - torch.roll(arr, 1, axis)


def _operator(epsilon, omega, shift):
  """Operator applied to ``(ww, 2, xx, yy, mm)`` fields (``_mode.py:22-50``)."""
  eps_yx = epsilon[[1, 0]][None, :, :, :, None]
  eps_z = epsilon[2][None, :, :, None]
  omega = omega[:, None, None, None, None]
  shift = shift[:, None, None, None, None]

  def apply(arr):
    a = (omega ** 2 * eps_yx - shift) * arr
    b = -_diff(arr[:, 0], -2, False) + _diff(arr[:, 1], -3, False)
    b = b / eps_z
    b = torch.stack([-_diff(b, -2, True), _diff(b, -3, True)], dim=1) * eps_yx
    c = _diff(arr[:, 0], -3, True) + _diff(arr[:, 1], -2, True)
    c = torch.stack([_diff(c, -3, False), _diff(c, -2, False)], dim=1)
    return a + b + c

  return apply


def _cross(arr, dx, dy, dz):
  fx, fy, fz = arr[:, 0], arr[:, 1], arr[:, 2]
  return torch.stack([dy(fz) - dz(fy), dz(fx) - dx(fz), dx(fy) - dy(fx)], dim=1)


def _curl(beta, arr, is_forward):
  """Curl of ``(ww, 3, xx, yy, mm)`` fields with d/dz = -i beta."""
  b = beta[:, None, None, :]
  return _cross(arr, dx=lambda x: _diff(x, -3, is_forward), dy=lambda x: _diff(x, -2, is_forward),
                dz=lambda x: -1j * b * x)


def _fullh(beta, x):
  hz = (_diff(x[:, 0], -3, True) + _diff(x[:, 1], -2, True)) / (1j * beta[:, None, None, :])
  return torch.stack([x[:, 0].to(hz.dtype), x[:, 1].to(hz.dtype), hz], dim=1)


def _e_from_h(beta, omega, epsilon, h):
  return _curl(beta, h, False) / (1j * omega[:, None, None, None, None] * epsilon[None, ..., None])


def _poynting(beta, omega, epsilon, x):
  h = _fullh(beta, x)
  e = _e_from_h(beta, omega, epsilon, h)
  return torch.real(torch.sum(e[:, 0] * h[:, 1] - e[:, 1] * h[:, 0], dim=(1, 2)))


def full_fields(beta, omega, epsilon, x):
  """(h, e, h2): the full H, the E it implies and the H that E implies
  (``_mode.py:79-90``); h2 == h for a solved mode."""
  beta = torch.as_tensor(beta)
  omega = torch.as_tensor(omega, device=beta.device)
  epsilon = torch.as_tensor(np.ascontiguousarray(epsilon), device=beta.device)
  x = torch.as_tensor(np.ascontiguousarray(x), device=beta.device)
  h = _fullh(beta, x)
  e = _e_from_h(beta, omega, epsilon, h)
  h2 = _curl(beta, e, True) / (-1j * omega[:, None, None, None, None])
  return h, e, h2


def _subspace_iteration(op, x, n, tol):
  """Subspace iteration (``_mode.py:99-142``): returns (w, x, err, iters),
  eigenvalues sorted in decreasing order."""
  ww, mm = x.shape[0], x.shape[-1]
  w = torch.zeros((ww, mm), dtype=x.dtype, device=x.device)
  err = torch.full((ww, mm), float("inf"), dtype=x.dtype, device=x.device)
  y = op(x)
  i = 0
  # The convergence test reads err on the host (a device sync): every
  # iteration for the first 8, then every 8th (iters may overshoot by < 8).
  while i < n and ((i > 8 and i % 8) or float(torch.max(err)) > tol):
    q, _ = torch.linalg.qr(torch.reshape(y, (ww, -1, mm)), mode="reduced")
    x = torch.reshape(q, x.shape)
    y = op(x)
    w = torch.sum(x * y, dim=(1, 2, 3))
    err = torch.linalg.vector_norm(torch.reshape(y - w[:, None, None, None, :] * x, (ww, -1, mm)),
                                   dim=1)
    i += 1
  inds = torch.flip(torch.argsort(w, dim=-1), dims=(-1,))
  x = torch.take_along_dim(x, inds[:, None, None, None, :], dim=-1)
  w = torch.take_along_dim(w, inds, dim=-1)
  err = torch.take_along_dim(err, inds, dim=-1)
  return w, x, err, i


def _subspace_iteration_hip(epsilon, omega, shift, x, n, tol):
  """The iteration of ``_subspace_iteration`` for the operator
  ``_operator(epsilon, omega, shift)``, run by the HIP kernel on the device
  of ``x`` (ww, 2, xx, yy, mm); same results, sorted the same way."""
  from . import _lib
  lib = _lib.load()
  ww, _, xx, yy, mm = x.shape
  dev = x.device
  f32 = dict(dtype=torch.float32, device=dev)
  eps = torch.as_tensor(epsilon, **f32).contiguous()
  om = torch.as_tensor(omega, **f32).reshape(ww).contiguous()
  sh = torch.as_tensor(shift, **f32).reshape(ww).contiguous()
  x = torch.as_tensor(x, **f32).contiguous().clone()
  nbytes = ctypes.c_size_t()
  _lib.check(lib.pjz_fdtd_mode_workspace_bytes(ww, xx, yy, mm, ctypes.byref(nbytes)), "mode")
  ws = torch.empty(max(nbytes.value // 4, 1), **f32)
  w = torch.empty((ww, mm), **f32)
  err = torch.empty((ww, mm), **f32)
  iters = torch.empty((ww,), dtype=torch.int32, device=dev)
  stream = torch.cuda.current_stream(dev).cuda_stream
  ptr = lambda t: ctypes.c_void_p(t.data_ptr())
  _lib.check(lib.pjz_fdtd_mode_solve(ww, xx, yy, mm, ptr(eps), ptr(om), ptr(sh), ptr(x), ptr(w),
                                     ptr(err), ptr(iters), int(n), float(tol), ptr(ws),
                                     nbytes.value, ctypes.c_void_p(stream)), "mode")
  inds = torch.flip(torch.argsort(w, dim=-1), dims=(-1,))
  x = torch.take_along_dim(x, inds[:, None, None, None, :], dim=-1)
  w = torch.take_along_dim(w, inds, dim=-1)
  err = torch.take_along_dim(err, inds, dim=-1)
  return w, x, err, int(iters.max())


HIP_MAX_MODES = 10


def _random(shape, device, seed=0):
  gen = torch.Generator(device="cpu").manual_seed(seed)
  return torch.randn(shape, generator=gen, dtype=torch.float32).to(device)


def mode(epsilon, omega, num_modes: int, init: Optional[torch.Tensor] = None,
         shift_iters: int = 10, max_iters: int = 100000, tol: float = 1e-4):
  """Waveguide modes of a cross-section; same arguments and results as pjz's
  ``mode()`` (``src/pjz/_mode.py:145-257``).

  ``epsilon`` is ``(3, xx, yy, zz)`` with exactly one of xx, yy, zz equal to
  1 (the propagation axis); ``omega`` is ``(ww,)``.  Returns
  ``(wavevector (ww, mm), excitation (ww, 2, xx, yy, zz, mm), err (ww, mm),
  iters)``, fundamental mode first, as torch tensors on ``epsilon``'s device.
  """
  device = epsilon.device if isinstance(epsilon, torch.Tensor) else torch.device("cpu")
  epsilon = torch.as_tensor(np.asarray(epsilon) if not isinstance(epsilon, torch.Tensor)
                            else epsilon, dtype=torch.float32, device=device)
  omega = torch.as_tensor(omega, dtype=torch.float32, device=device)
  if omega.ndim == 0:
    omega = omega[None]
  shape = tuple(epsilon.shape)
  if 1 not in shape[1:]:
    raise ValueError(f"Expected exactly one of the spatial dimensions of ``epsilon`` to be "
                     f"singular, instead got ``epsilon.shape == {shape}``.")
  prop_axis = "xyz"[shape[1:].index(1)]
  ww = omega.shape[0]
  plane = tuple(s for i, s in enumerate(shape[1:]) if i != "xyz".index(prop_axis))
  mode_shape = (ww, 2) + plane + (num_modes,)

  if init is None:
    init = _random(mode_shape, device)
  else:
    init = torch.as_tensor(init, dtype=torch.float32, device=device)
    init = torch.squeeze(init, "xyz".index(prop_axis) + 2)
    if prop_axis == "y":
      init = torch.flip(torch.swapaxes(init, 2, 3), dims=(1, 2))
    elif prop_axis == "z":
      init = init * torch.tensor([1.0, -1.0], device=device)[None, :, None, None, None]
    init = torch.flip(init, dims=(1,))  # excitation-to-field flip

  # Propagate-along-z form.
  if prop_axis == "x":
    epsilon = epsilon[[1, 2, 0]]
  elif prop_axis == "y":
    epsilon = torch.flip(torch.swapaxes(epsilon[[2, 0, 1]], 1, 3), dims=(1,))
    mode_shape = tuple(mode_shape[i] for i in (0, 1, 3, 2, 4))
  epsilon = torch.squeeze(epsilon)
  init = torch.reshape(init, mode_shape)

  if epsilon.is_cuda and num_modes <= HIP_MAX_MODES:
    shift, _, _, _ = _subspace_iteration_hip(epsilon, omega, torch.zeros_like(omega),
                                             _random(mode_shape[:-1] + (1,), device), shift_iters,
                                             tol)
    w, x, err, iters = _subspace_iteration_hip(epsilon, omega, shift[:, 0], init, max_iters, tol)
  else:
    shift, _, _, _ = _subspace_iteration(_operator(epsilon, omega, torch.zeros_like(omega)),
                                         _random(mode_shape[:-1] + (1,), device), shift_iters, tol)
    w, x, err, iters = _subspace_iteration(_operator(epsilon, omega, shift[:, 0]), init,
                                           max_iters, tol)
  wavevector = torch.sqrt(w + shift)
  x = x / torch.sqrt(_poynting(wavevector, omega, epsilon, x))[:, None, None, None, :]
  exc = torch.flip(x, dims=(1,))  # field-to-excitation flip
  if prop_axis == "y":
    exc = torch.swapaxes(torch.flip(exc, dims=(1, 2)), 2, 3)
  elif prop_axis == "z":
    exc = exc * torch.tensor([1.0, -1.0], device=device)[None, :, None, None, None]
  exc = torch.unsqueeze(exc, "xyz".index(prop_axis) + 2)
  return (wavevector.to(torch.float32), exc.to(torch.float32), err.to(torch.float32), iters)
