"""numpy restatement of pjz's scatter() -- TEST INFRASTRUCTURE ONLY.

Follows ``src/pjz/_field.py:293-398`` (overlaps, amplitudes, S-values and the
reciprocity gradient) on top of ``field_ref.field``.  Pinning: the 2x2
amplitude solve is checked against a closed form (a pure forward wave
exp(i beta x) gives amplitudes (1, 0)); the rest follows the cited lines
(the reference has no scatter() test).
"""

import numpy as np

from . import field_ref
from . import helpers as H


def amplitudes(beta, vals, x):  # _field.py:293-302
  bx = beta[:, None] * np.asarray(x, np.float64)[None, :]
  a = np.stack([np.exp(-1j * bx), np.exp(1j * bx)], axis=1)
  return np.einsum("...ji,...j->...i", np.linalg.pinv(a), vals)


def overlap(mode, beta, pos, is_fwd, output):  # _field.py:305-338
  beta = np.asarray(beta, np.complex128)
  if is_fwd is None:
    x, sample_at, beta = np.array([0, 0]), (pos, pos), beta * 0
  elif is_fwd:
    x, sample_at = np.array([1, 2]), (pos + 1, pos + 2)
  else:
    x, sample_at, beta = np.array([-2, -1]), (pos - 2, pos - 1), -beta
  axis = H.prop_axis(mode.shape)
  vals = np.stack([np.sum(mode * field_ref._transverse_slice(output, p, axis),
                          axis=(-4, -3, -2, -1)) for p in sample_at], axis=-1)
  return amplitudes(beta, vals, x)


def scatter(epsilon, omega, modes, betas, pos, is_fwd, p):
  """(svals, grads, fields) as _scatter_impl builds them (_field.py:346-384)."""
  fields = [field_ref.field(epsilon, np.mean(m, axis=0), omega, q, p)
            for m, q in zip(modes, pos)]
  amps = [np.ones_like(b, dtype=np.complex128) if fwd is None else overlap(m, b, q, fwd, f)[:, 0]
          for f, m, b, q, fwd in zip(fields, modes, betas, pos, is_fwd)]
  svals = [[overlap(m, b, q, fwd, f)[:, 1] / a for m, b, q, fwd in zip(modes, betas, pos, is_fwd)]
           for a, f in zip(amps, fields)]
  grads = [[fi * fj / a[:, None, None, None, None] for fj in fields]
           for a, fi in zip(amps, fields)]
  return svals, grads, fields


def scatter_vjp(grads, g):  # _field.py:393-398
  return sum(sum(np.sum(np.real(gij[:, None, None, None, None] * gradij), axis=0)
                 for gradij, gij in zip(gradi, gi))
             for gradi, gi in zip(grads, g))
