"""Scattering parameters between ports: ``scatter()``.

Mirrors ``src/pjz/_field.py:293-442``: one ``field()`` simulation per port
(the source is the port mode averaged over frequency), mode overlaps on two
planes next to each port split into forward/backward amplitudes, S-values
normalised by the injected amplitude, and a custom backward pass that
returns pjz's reciprocity gradient  sum_ij Re(g_ij * f_i * f_j / a_i)
(``_field.py:380-382,393-398``) instead of differentiating through the
simulations.  ``g_ij`` there is jax's cotangent dL/dRe - i dL/dIm; torch's is
its conjugate, which the backward conjugates back.  The per-port
simulations run as one batch on the GPU (:func:`pjz_amd._field.fields`).
"""

from typing import Any, Sequence

import torch

from ._field import SimParams, fields as port_fields, prop_axis, transverse_slice


def amplitudes(beta, vals, x):
  """Forward/backward amplitudes from two overlap samples.

  Mirrors ``_amplitudes``, ``_field.py:293-302``: solves the 2x2 system
  [exp(-i beta x); exp(i beta x)] via its pseudo-inverse, contracted as
  ``einsum("...ji,...j->...i", pinv(a), vals)``.
  """
  bx = beta[:, None] * x[None, :].to(device=beta.device, dtype=beta.dtype)
  a = torch.stack([torch.exp(-1j * bx), torch.exp(1j * bx)], dim=1)
  return torch.einsum("...ji,...j->...i", torch.linalg.pinv(a), vals)


def overlap(mode, beta, pos, is_fwd, output):
  """(ww, 2) amplitudes of ``mode`` in ``output`` next to port ``pos``.

  Mirrors ``_overlap``, ``_field.py:305-338``: samples at (pos, pos) with
  beta = 0 when ``is_fwd`` is None, at (pos+1, pos+2) for a forward port and
  at (pos-2, pos-1) for a backward one (beta negated).
  """
  beta = beta.to(torch.complex128)
  if is_fwd is None:
    x, sample_at, beta = torch.tensor([0.0, 0.0]), (pos, pos), beta * 0
  elif is_fwd:
    x, sample_at = torch.tensor([1.0, 2.0]), (pos + 1, pos + 2)
  else:
    x, sample_at, beta = torch.tensor([-2.0, -1.0]), (pos - 2, pos - 1), -beta
  axis = prop_axis(mode.shape)
  m = mode.to(torch.complex128)
  vals = torch.stack(
      [torch.sum(m * transverse_slice(output, p, axis).to(torch.complex128), dim=(-4, -3, -2, -1))
       for p in sample_at], dim=-1)
  return amplitudes(beta.to(vals.device), vals, x.to(vals.device))


def _impl(epsilon, omega, modes, betas, pos, is_fwd, sim_params):
  # One simulation per port (source = the port mode averaged over frequency),
  # all ports advanced together by the batched kernel.
  fields = port_fields(epsilon, [torch.mean(m, dim=0) for m in modes], omega, pos, sim_params)
  amps = [torch.ones_like(b, dtype=torch.complex128) if fwd is None
          else overlap(m, b, p, fwd, f)[:, 0]
          for f, m, b, p, fwd in zip(fields, modes, betas, pos, is_fwd)]
  svals = [[overlap(m, b, p, fwd, f)[:, 1] / a
            for m, b, p, fwd in zip(modes, betas, pos, is_fwd)]
           for a, f in zip(amps, fields)]
  return svals, amps, fields


class _Scatter(torch.autograd.Function):

  @staticmethod
  def forward(ctx, epsilon, omega, modes, betas, pos, is_fwd, sim_params):
    with torch.no_grad():
      svals, amps, fields = _impl(epsilon, omega, modes, betas, pos, is_fwd, sim_params)
    ctx.amps, ctx.fields = amps, fields
    ctx.eps_dtype, ctx.eps_device = epsilon.dtype, epsilon.device
    n = len(modes)
    return torch.stack([torch.stack(row) for row in svals]).reshape(n, n, -1)

  @staticmethod
  def backward(ctx, g):
    fields, amps = ctx.fields, ctx.amps
    grad = None
    for i, (fi, ai) in enumerate(zip(fields, amps)):
      for j, fj in enumerate(fields):
        gradij = fi * fj / ai[:, None, None, None, None].to(fi.dtype)
        # torch hands complex outputs the cotangent dL/dRe + i dL/dIm, the
        # conjugate of the one jax passes to pjz's _scatter_bwd; conjugating it
        # gives the same gradient as the reference for the same loss.
        gij = torch.conj(g[i, j])[:, None, None, None, None].to(gradij.dtype)
        term = torch.sum(torch.real(gij * gradij), dim=0)
        grad = term if grad is None else grad + term
    return (grad.to(device=ctx.eps_device, dtype=ctx.eps_dtype), None, None, None, None, None,
            None)


def scatter(epsilon, omega, modes: Sequence[torch.Tensor], betas: Sequence[torch.Tensor],
            pos: Sequence[int], is_fwd: Sequence[Any], sim_params: SimParams):
  """Scattering values ``svals[i][j]`` (ww,) from port i to port j.

  Same arguments and result as ``src/pjz/_field.py:403-439``; differentiable
  with respect to ``epsilon`` through pjz's reciprocity gradient.
  """
  modes = tuple(modes)
  dev = epsilon.device if isinstance(epsilon, torch.Tensor) else None
  betas = tuple(torch.as_tensor(b, device=dev) for b in betas)
  out = _Scatter.apply(epsilon, omega, modes, betas, tuple(pos), tuple(is_fwd), sim_params)
  n = len(modes)
  return [[out[i, j] for j in range(n)] for i in range(n)]
