"""Boundary, waveform and frequency helpers on the fdtdz() call path.

Host-side glue that turns pjz's simulation parameters into fdtdz() inputs and
snapshots back into complex fields.  Each function keeps the name the
reference tests use (tests/test_boundaries.py, test_waveforms.py,
test_frequencies.py) and the formula of the pjz function it mirrors.  Results
are float32 numpy arrays (complex64 for complex results), the dtypes pjz's
jax.numpy code produces, except ``sampling_interval`` (an int).
"""

import numpy as np


def zz(pml_widths, use_reduced_precision):
  """Height of the simulation domain inside the PML budget.

  Mirrors ``_zz``, ``src/pjz/_field.py:56-58``.
  """
  return (128 if use_reduced_precision else 64) - sum(pml_widths)


def pad_zz(epsilon_zz, pml_widths, use_reduced_precision):
  """(bottom, top) z padding that centres epsilon in the domain.

  Mirrors ``_pad_zz``, ``src/pjz/_field.py:61-66``.
  """
  total = zz(pml_widths, use_reduced_precision)
  if epsilon_zz > total:
    raise ValueError(f"epsilon has {epsilon_zz} z cells, the domain only {total}")
  bot = (total - epsilon_zz) // 2
  return bot, total - epsilon_zz - bot


def _absorption_profiles(numcells, width, smoothness):
  centre = (numcells - 1) / 2
  pos = np.arange(numcells, dtype=np.float64)[None, :] + np.array([[0.0], [0.5]])
  pos = np.clip(np.abs(pos - centre) - centre + width, 0.0, None)
  return smoothness * pos ** 2


def absorption_mask(xx, yy, width, smoothness):
  """(3, xx, yy) absorber conductivity, quadratic over ``width`` edge cells.

  Mirrors ``_absorption_mask``, ``src/pjz/_field.py:69-90``: component 0 uses
  the x profile at integer samples and the y profile at half samples,
  component 1 the reverse, component 2 half samples in both.
  """
  x = _absorption_profiles(xx, width, smoothness)
  y = _absorption_profiles(yy, width, smoothness)
  cross = lambda a, b: np.maximum(a[:, None], b[None, :])
  return np.stack([cross(x[0], y[1]), cross(x[1], y[0]),
                   cross(x[1], y[1])]).astype(np.float32)


def pml_sigma(pml_widths, zz, ln_R, m):
  """(zz, 2) z-PML conductivity; column 0 at z = k, column 1 at z = k + 1/2.

  Mirrors ``_pml_sigma``, ``src/pjz/_field.py:93-106``.
  """
  lo, hi = pml_widths
  z = np.arange(zz, dtype=np.float64)[None, :] + np.array([[0.0], [0.5]])
  dlo = (lo - z) / lo if lo != 0 else np.zeros_like(z)
  dhi = (z + 0.5 - zz + hi) / hi if hi != 0 else np.zeros_like(z)
  d = np.max(np.clip(np.stack([dlo, dhi], axis=-1), 0.0, None), axis=-1)
  return ((m + 1) * ln_R * d ** m).T.astype(np.float32)


def ramped_sin(omega, width, delay, dt, tt):
  """Complex sinusoid with a tanh ramp, averaged over ``omega``; (tt,) complex.

  Mirrors ``_ramped_sin``, ``src/pjz/_field.py:109-113``.  Its imaginary part
  tends to sin(omega * dt * t).
  """
  omega = np.atleast_1d(np.asarray(omega, dtype=np.float64))
  t = omega[:, None] * dt * np.arange(tt)
  w = ((1 + np.tanh(t / width - delay)) / 2) * np.exp(1j * t)
  return np.mean(w, axis=0).astype(np.complex64)


def sampling_interval(omega_min, omega_max, omega_n, dt):
  """Snapshot interval that keeps ``omega_n`` components separable.

  Mirrors ``_sampling_interval``, ``src/pjz/_field.py:116-139``.
  """
  period = 4 * np.pi / (omega_max + omega_min) / dt
  if omega_n == 1:
    return int(round(period / 4))
  cutoff = np.pi * (omega_n - 1) / omega_n / (omega_max - omega_min) / dt
  m = np.floor((cutoff - period / 4) / (period / 2))
  return int(round(period / 4 + m * period / 2))


def output_phases(omega, steps, dt):
  """(2 ww, nn) matrix [cos(theta); -sin(theta)], theta = omega dt step.

  Mirrors ``_output_phases``, ``src/pjz/_field.py:142-150``.
  """
  omega = np.atleast_1d(np.asarray(omega, dtype=np.float64))
  theta = omega[:, None] * dt * np.asarray(steps, dtype=np.float64)[None, :]
  return np.concatenate([np.cos(theta), -np.sin(theta)], axis=0)


def projection_matrix(omega, steps, dt):
  """pinv(phases.T): (2 ww, nn) float64 map from snapshots to [Re; Im]."""
  return np.linalg.pinv(output_phases(omega, steps, dt).T)


def frequency_components(fields, steps, omega, dt):
  """Complex amplitudes (ww, ...) of real snapshots ``fields`` (nn, ...).

  Mirrors the projection at ``src/pjz/_field.py:271-279``: a signal
  sin(omega dt s) maps to -i, cos(omega dt s) to 1.
  """
  omega = np.atleast_1d(np.asarray(omega, dtype=np.float64))
  out = np.einsum("ij,j...->i...", projection_matrix(omega, steps, dt),
                  np.asarray(fields, dtype=np.float64))
  ww = omega.shape[0]
  return (out[:ww] + 1j * out[ww:]).astype(np.complex64)


def source_amplitude(waveform, omega, dt):
  """Complex amplitude (ww, k) of a (tt, k) source waveform.

  Waveform sample t drives the E update of step t, which is centred at time
  (t - 1/2) dt relative to the E snapshot of step s at time s dt; so a sine
  sampled at t - 1/2 has amplitude -i (tests/test_frequencies.py:32-38).
  Least-squares fit over all samples with the phase convention of
  ``output_phases``.
  """
  waveform = np.asarray(waveform, dtype=np.float64)
  return frequency_components(waveform, np.arange(waveform.shape[0]) - 0.5, omega, dt)
