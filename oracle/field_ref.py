"""numpy restatement of pjz's field() glue -- TEST INFRASTRUCTURE ONLY.

Follows ``src/pjz/_field.py:171-279`` step by step (z padding, waveform,
output schedule, source scaling / padding / odd-position shift / quadratures,
boundaries, kernel call, pinv projection) with the oracle helpers and the
FP64 C stepper.  Pinning: each helper is pinned in tests/test_oracle_golden.py
and the stepper in tests/test_oracle_*.py; the glue itself follows the cited
lines (parity unpinned beyond code reading: the reference has no field()
test).
"""

import numpy as np

from . import c_oracle
from . import helpers as H


def _transverse_slice(arr, pos, axis):  # _field.py:153-163
  a = "xyz".find(axis)
  comps = [i for i in range(3) if i != a]
  arr = arr[..., comps, :, :, :]
  idx = [slice(None)] * arr.ndim
  idx[arr.ndim - 3 + a] = slice(pos, pos + 1)
  return arr[tuple(idx)]


def inputs(epsilon, source, omega, source_pos, p):
  """fdtdz() keyword arguments exactly as field() builds them."""
  omega = np.atleast_1d(np.asarray(omega, np.float64))
  ww = omega.shape[0]
  pad_zz = H.pad_zz(epsilon.shape[3], p.pml_widths, p.use_reduced_precision)
  ap = p.absorption_padding
  padding = ((ap, ap), (ap, ap), pad_zz)
  rsin = H.ramped_sin(omega, p.source_ramp, p.source_delay, p.dt, p.tt)
  if source.shape[3] == 1:
    wave = np.stack([np.imag(rsin), np.real(rsin)], axis=-1)
  else:
    wave = np.pad(np.imag(rsin)[:, None], ((0, 0), (0, 1)))
  interval = H.sampling_interval(p.omega_range[0], p.omega_range[1], ww, p.dt)
  output_steps = (p.tt - 2 * interval * ww - 1, p.tt, interval)
  src = source / _transverse_slice(epsilon, source_pos, H.prop_axis(source.shape))
  pos = int(source_pos)
  for i in range(3):
    if src.shape[i + 1] == 1:
      pos += padding[i][0]
    else:
      src = np.pad(src, (i + 1) * ((0, 0),) + (padding[i],) + (2 - i) * ((0, 0),))
  if (src.shape[1] == 1 or src.shape[2] == 1) and pos % 2 == 1:
    pos += 1
    wave = np.flip(wave, axis=1)
  elif src.shape[3] == 1:
    src = np.stack([np.imag(src), np.real(src)])
  src = np.real(src)
  mask = H.absorption_mask(epsilon.shape[1] + 2 * ap, epsilon.shape[2] + 2 * ap, ap,
                           p.absorption_coeff)
  ZZ = H.zz(p.pml_widths, p.use_reduced_precision)
  if p.use_z_as_batch:
    kappa = np.full((ZZ, 2), np.inf)
    sigma = np.zeros((ZZ, 2))
    alpha = np.zeros((ZZ, 2))
  else:
    sigma = H.pml_sigma(p.pml_widths, ZZ, p.pml_sigma_lnr, p.pml_sigma_m)
    kappa = np.ones_like(sigma)
    alpha = p.pml_alpha_coeff * np.ones_like(sigma)
  kw = dict(epsilon=epsilon, dt=p.dt, source_field=src, source_waveform=wave,
            source_position=pos, absorption_mask=mask, pml_kappa=kappa, pml_sigma=sigma,
            pml_alpha=alpha, pml_widths=p.pml_widths, output_steps=output_steps,
            use_reduced_precision=p.use_reduced_precision, launch_params=None,
            offset=(padding[0][0], padding[1][0], padding[2][0]))
  return kw


def field(epsilon, source, omega, source_pos, p):
  """(ww, 3, xx, yy, zz) complex128 fields, _field.py:171-279 on the oracle."""
  kw = inputs(np.asarray(epsilon, np.float64), np.asarray(source), omega, source_pos, p)
  fields = c_oracle.fdtdz(**kw)
  return H.frequency_components(fields, np.arange(*kw["output_steps"]), omega, p.dt)
