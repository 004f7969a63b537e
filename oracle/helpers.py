"""numpy FP64 restatements of pjz's host-side helpers -- TEST INFRASTRUCTURE ONLY.

Each function restates the formula of the pjz function cited next to it.  They
check the product's own helpers (``pjz_amd.boundaries`` etc.) and build
inputs for the stepper oracle.

Pinning:
  absorption_mask    literal rows of tests/test_boundaries.py:8-24
  pml_sigma          literal 10x2 table of tests/test_boundaries.py:27-41
  ramped_sin         asymptote to sin(w*dt*t), tests/test_waveforms.py:8-14
  sampling_interval  near-orthogonality, tests/test_frequencies.py:8-18
  output_phases /
  frequency_components  sin->-i, cos->1, tests/test_frequencies.py:21-29, and
                     exact recovery of synthetic multi-tone signals
  source_amplitude   half-step-shifted sine -> -i, tests/test_frequencies.py:32-38
  pad_zz / zz        closed-form integer identities (sum of pads, centring)
"""

import numpy as np


def zz(pml_widths, use_reduced_precision):
  """Kernel z budget.  Follows ``src/pjz/_field.py:56-58``."""
  return (128 if use_reduced_precision else 64) - sum(pml_widths)


def pad_zz(epsilon_zz, pml_widths, use_reduced_precision):
  """(bot, top) z padding.  Follows ``src/pjz/_field.py:61-66``."""
  total = zz(pml_widths, use_reduced_precision)
  bot = (total - epsilon_zz) // 2
  return bot, total - epsilon_zz - bot


def _profiles(n, width, smoothness):
  """Follows ``src/pjz/_field.py:69-76``: rows for sample offsets 0 and 1/2."""
  centre = (n - 1) / 2.0
  pos = np.arange(n)[None, :] + np.array([[0.0], [0.5]])
  d = np.clip(np.abs(pos - centre) - centre + width, 0.0, None)
  return smoothness * d ** 2


def absorption_mask(xx, yy, width, smoothness):
  """(3, xx, yy) absorber.  Follows ``src/pjz/_field.py:79-90``."""
  px = _profiles(xx, width, smoothness)
  py = _profiles(yy, width, smoothness)
  cross = lambda a, b: np.maximum(a[:, None], b[None, :])
  return np.stack([cross(px[0], py[1]), cross(px[1], py[0]),
                   cross(px[1], py[1])])


def pml_sigma(pml_widths, zz_, ln_R, m):
  """(zz, 2) PML conductivity.  Follows ``src/pjz/_field.py:93-106``."""
  lo, hi = pml_widths
  z = np.arange(zz_)[None, :] + np.array([[0.0], [0.5]])
  dlo = (lo - z) / lo if lo != 0 else np.zeros_like(z)
  dhi = (z + 0.5 - zz_ + hi) / hi if hi != 0 else np.zeros_like(z)
  d = np.max(np.clip(np.stack([dlo, dhi], -1), 0.0, None), axis=-1)
  return ((m + 1) * ln_R * d ** m).T


def ramped_sin(omega, width, delay, dt, tt):
  """Complex ramped sinusoid averaged over omega.  ``src/pjz/_field.py:109-113``."""
  omega = np.atleast_1d(np.asarray(omega, dtype=np.float64))
  t = omega[:, None] * dt * np.arange(tt)
  w = ((1 + np.tanh(t / width - delay)) / 2) * np.exp(1j * t)
  return np.mean(w, axis=0)


def sampling_interval(omega_min, omega_max, omega_n, dt):
  """Snapshot interval.  Follows ``src/pjz/_field.py:116-139``."""
  period = 4 * np.pi / (omega_max + omega_min) / dt
  if omega_n == 1:
    return int(round(period / 4))
  cutoff = np.pi * (omega_n - 1) / omega_n / (omega_max - omega_min) / dt
  m = np.floor((cutoff - period / 4) / (period / 2))
  return int(round(period / 4 + m * period / 2))


def output_phases(omega, steps, dt):
  """(2 ww, nn) phase matrix [cos; -sin].  Follows ``src/pjz/_field.py:142-150``."""
  omega = np.atleast_1d(np.asarray(omega, dtype=np.float64))
  theta = omega[:, None] * dt * np.asarray(steps, dtype=np.float64)[None, :]
  return np.concatenate([np.cos(theta), -np.sin(theta)], axis=0)


def frequency_components(fields, steps, omega, dt):
  """Complex amplitudes from snapshots.  Follows ``src/pjz/_field.py:271-279``."""
  omega = np.atleast_1d(np.asarray(omega, dtype=np.float64))
  ph = output_phases(omega, steps, dt)
  out = np.einsum("ij,j...->i...", np.linalg.pinv(ph.T), np.asarray(fields))
  ww = omega.shape[0]
  return out[:ww] + 1j * out[ww:]


def source_amplitude(waveform, omega, dt):
  """Complex amplitude of a (tt, k) source waveform at each omega.

  Waveform sample t is taken to sit at time (t - 1/2) * dt, the convention
  pinned by ``tests/test_frequencies.py:32-38`` (a sine sampled at t - 1/2
  yields -i).  Least-squares fit over all samples, same phase convention as
  ``output_phases``.  Returns (ww, k) complex.
  """
  waveform = np.asarray(waveform, dtype=np.float64)
  t = np.arange(waveform.shape[0]) - 0.5
  return frequency_components(waveform, t, omega, dt)


def prop_axis(shape):
  """Axis with extent 1 among the last three.  ``src/pjz/_field.py:282-290``."""
  last = tuple(shape[-3:])
  if last.count(1) != 1:
    raise ValueError("mode.shape[-3:] must contain exactly one value of 1")
  return "xyz"[last.index(1)]
