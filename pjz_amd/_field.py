"""Time-harmonic fields from the FDTD kernel: ``SimParams`` and ``field()``.

Mirrors pjz's simulation driver (``src/pjz/_field.py:12-279``) on torch
tensors: the same parameters and defaults, the same input glue (z padding,
waveform, output schedule, source scaling and placement, boundaries), one call
to :func:`pjz_amd.fdtdz`, and the projection of the snapshots onto complex
frequency components, done on the GPU.
"""

from typing import Any, NamedTuple, Tuple

import numpy as np
import torch

from . import boundaries as B
from .fdtdz_torch import fdtdz, fdtdz_batch


class SimParams(NamedTuple):
  """Simulation parameters; fields and defaults of ``src/pjz/_field.py:12-53``.

  Attributes:
    omega_range: (omega_min, omega_max) of the frequencies extracted from one
      simulation.
    tt: number of FDTD updates.
    dt: time step.
    source_ramp, source_delay: tanh ramp width and delay of the source, in
      radians of source phase.
    absorption_padding: x-y cells added on each side for the absorber.
    absorption_coeff: strength of the quadratic absorber.
    pml_widths: (bottom, top) z-PML cells.
    pml_alpha_coeff, pml_sigma_lnr, pml_sigma_m: CPML alpha and sigma profile.
    use_z_as_batch: decouple z slices (kappa = inf, sigma = alpha = 0).
    use_reduced_precision: selects the 128-cell z budget (storage stays fp32).
    launch_params: passed through to fdtdz().
  """
  omega_range: Tuple[float, float]
  tt: int
  dt: float = 0.5
  source_ramp: float = 4.0
  source_delay: float = 4.0
  absorption_padding: int = 50
  absorption_coeff: float = 1e-4
  pml_widths: Tuple[int, int] = (16, 16)
  pml_alpha_coeff: float = 0.0
  pml_sigma_lnr: float = 0.5
  pml_sigma_m: float = 1.3
  use_z_as_batch: bool = False
  use_reduced_precision: bool = True
  launch_params: Any = None


def prop_axis(shape):
  """'x', 'y' or 'z': the singular axis among the last three of ``shape``.

  Mirrors ``_prop_axis``, ``src/pjz/_field.py:282-290``.
  """
  last = tuple(shape[-3:])
  if last.count(1) != 1:
    raise ValueError(f"``mode.shape[-3:]`` must contain exactly one value of ``1``, "
                     f"instead got ``mode.shape == {tuple(shape)}``.")
  return "xyz"[last.index(1)]


def transverse_slice(arr, pos, axis):
  """Transverse components of ``arr`` on the plane ``pos`` normal to ``axis``.

  Mirrors ``_transverse_slice``, ``src/pjz/_field.py:153-163``: drops the
  component along ``axis`` and keeps a width-1 slab at ``pos``.
  """
  a = "xyz".find(axis)
  comps = [i for i in range(3) if i != a]
  arr = arr[..., comps, :, :, :]
  dim = arr.ndim - 3 + a
  return arr.narrow(dim, int(pos), 1)


def _as_tensor(x, device, dtype=None):
  if isinstance(x, torch.Tensor):
    return x.to(device=device, dtype=dtype) if dtype is not None else x.to(device)
  x = torch.as_tensor(np.asarray(x))
  return x.to(device=device, dtype=dtype) if dtype is not None else x.to(device)


def simulation_inputs(epsilon, source, omega, source_pos, sim_params):
  """fdtdz() keyword arguments for pjz's field() (``_field.py:194-269``).

  Returns (kwargs, output_steps).  Tensors are on ``epsilon``'s device.
  """
  p = sim_params
  dev = epsilon.device
  omega_np = np.atleast_1d(np.asarray(
      omega.detach().cpu().numpy() if isinstance(omega, torch.Tensor) else omega, np.float64))
  ww = omega_np.shape[0]
  pad_zz = B.pad_zz(epsilon.shape[3], p.pml_widths, p.use_reduced_precision)
  ap = p.absorption_padding
  padding = ((ap, ap), (ap, ap), pad_zz)

  rsin = B.ramped_sin(omega_np, p.source_ramp, p.source_delay, p.dt, p.tt)
  if source.shape[3] == 1:  # z-plane source: two quadratures
    wave = np.stack([np.imag(rsin), np.real(rsin)], axis=-1)
  else:
    wave = np.stack([np.imag(rsin), np.zeros_like(np.imag(rsin))], axis=-1)

  interval = B.sampling_interval(p.omega_range[0], p.omega_range[1], ww, p.dt)
  output_steps = (p.tt - 2 * interval * ww - 1, p.tt, interval)

  # Source scaled by epsilon on its plane (``_source``, _field.py:166-167).
  axis = prop_axis(source.shape)
  src = source / transverse_slice(epsilon, source_pos, axis).to(source.dtype)
  pos = int(source_pos)
  for i in range(3):
    if src.shape[i + 1] == 1:
      pos += padding[i][0]
    else:
      pad = [0, 0] * 3
      # torch pads the last dimension first
      j = 2 - i
      pad[2 * j], pad[2 * j + 1] = padding[i]
      src = torch.nn.functional.pad(src, pad)
  if (src.shape[1] == 1 or src.shape[2] == 1) and pos % 2 == 1:
    pos += 1
    wave = wave[:, ::-1]
  elif src.shape[3] == 1:
    src = torch.stack([torch.imag(src), torch.real(src)]) if torch.is_complex(src) else \
        torch.stack([torch.zeros_like(src), src])
  if torch.is_complex(src):
    src = torch.real(src)

  XX = epsilon.shape[1] + 2 * ap
  YY = epsilon.shape[2] + 2 * ap
  ZZ = B.zz(p.pml_widths, p.use_reduced_precision)
  mask = B.absorption_mask(XX, YY, ap, p.absorption_coeff)
  if p.use_z_as_batch:
    kappa = np.full((ZZ, 2), np.inf, np.float32)
    sigma = np.zeros((ZZ, 2), np.float32)
    alpha = np.zeros((ZZ, 2), np.float32)
  else:
    sigma = B.pml_sigma(p.pml_widths, ZZ, p.pml_sigma_lnr, p.pml_sigma_m)
    kappa = np.ones_like(sigma)
    alpha = p.pml_alpha_coeff * np.ones_like(sigma)
  f32 = lambda a: _as_tensor(np.ascontiguousarray(a), dev, torch.float32)
  kw = dict(epsilon=epsilon.to(torch.float32), dt=p.dt,
            source_field=src.to(torch.float32).contiguous(),
            source_waveform=f32(wave), source_position=pos, absorption_mask=f32(mask),
            pml_kappa=f32(kappa), pml_sigma=f32(sigma), pml_alpha=f32(alpha),
            pml_widths=tuple(p.pml_widths), output_steps=output_steps,
            use_reduced_precision=p.use_reduced_precision, launch_params=p.launch_params,
            offset=(padding[0][0], padding[1][0], padding[2][0]))
  return kw, output_steps


def project(fields, omega, output_steps, dt):
  """Snapshots (nn, ...) -> complex amplitudes (ww, ...) on the GPU.

  ``einsum(pinv(phases.T), fields)`` as at ``src/pjz/_field.py:271-279``; the
  (2 ww, nn) pseudo-inverse is formed in float64 on the host.
  """
  omega_np = np.atleast_1d(np.asarray(
      omega.detach().cpu().numpy() if isinstance(omega, torch.Tensor) else omega, np.float64))
  ww = omega_np.shape[0]
  steps = np.arange(*output_steps)
  P = torch.as_tensor(B.projection_matrix(omega_np, steps, dt), dtype=torch.float32,
                      device=fields.device)
  flat = fields.reshape(fields.shape[0], -1)
  out = (P @ flat).reshape((2 * ww,) + tuple(fields.shape[1:]))
  return torch.complex(out[:ww], out[ww:])


def field(epsilon, source, omega, source_pos, sim_params):
  """Time-harmonic solution of Maxwell's equations (``_field.py:171-279``).

  Args:
    epsilon: (3, xx, yy, zz) permittivity (torch tensor on the GPU, or numpy).
    source: (2, 1, yy, zz), (2, xx, 1, zz) or (2, xx, yy, 1) mode profile.
    omega: (ww,) angular frequencies.
    source_pos: position of the source along its propagation axis.
    sim_params: SimParams.

  Returns:
    (ww, 3, xx, yy, zz) complex64 fields at ``omega``.
  """
  dev = epsilon.device if isinstance(epsilon, torch.Tensor) and epsilon.is_cuda else \
      torch.device("cuda", torch.cuda.current_device()) if torch.cuda.is_available() else None
  if dev is None:
    raise RuntimeError("field() needs a HIP GPU: torch.cuda.is_available() is False")
  epsilon = _as_tensor(epsilon, dev, torch.float32)
  source = _as_tensor(source, dev)
  kw, output_steps = simulation_inputs(epsilon, source, omega, source_pos, sim_params)
  fields = fdtdz(**kw)
  return project(fields, omega, output_steps, sim_params.dt)


def _batchable(zz):
  """Whether the batched kernels take a padded domain height of zz."""
  return zz <= 64 or (zz <= 128 and zz % 2 == 0)


def fields(epsilon, sources, omega, source_positions, sim_params):
  """``field()`` for several sources on one structure (pjz's per-port loop
  in ``scatter()``, ``_field.py:354-356``), advanced together by one batched
  simulation (:func:`pjz_amd.fdtdz_batch`) when the padded domain height
  allows, otherwise one simulation per source.  Returns a list of
  (ww, 3, xx, yy, zz) complex64 fields, equal to ``field()`` per source.
  """
  dev = epsilon.device if isinstance(epsilon, torch.Tensor) and epsilon.is_cuda else \
      torch.device("cuda", torch.cuda.current_device()) if torch.cuda.is_available() else None
  if dev is None:
    raise RuntimeError("fields() needs a HIP GPU: torch.cuda.is_available() is False")
  epsilon = _as_tensor(epsilon, dev, torch.float32)
  runs = [simulation_inputs(epsilon, _as_tensor(src, dev), omega, pos, sim_params)
          for src, pos in zip(sources, source_positions)]
  if len(runs) > 1 and all(_batchable(kw["pml_sigma"].shape[0]) for kw, _ in runs):
    outs = fdtdz_batch([kw for kw, _ in runs])
  else:
    outs = [fdtdz(**kw) for kw, _ in runs]
  return [project(f, omega, steps, sim_params.dt) for f, (_, steps) in zip(outs, runs)]
