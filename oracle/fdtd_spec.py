"""FP64 stepper of the fdtdz() Yee update -- TEST INFRASTRUCTURE ONLY.

The reference kernel ``fdtdz_jax.fdtdz`` (PyPI ``fdtdz>=1.1.3``, pinned at
``setup.py:26``, called at ``src/pjz/_field.py:254-269``) is not vendored under
the reference tree, so its update equations cannot be read.  This module states
the update equations this project implements, following the call contract that
pjz's own code fixes (array shapes and roles at ``_field.py:189-191,199-269``;
Yee staggering at ``_epsilon.py:13-24,151-152`` and ``_field.py:69-106``) and
the published methods they name (Yee 1966; Roden & Gedney 2000 CPML).

Update for step n = 0, 1, ... (unit cells, mu = 1, c = 1):

  H half step  H -= dt * curl_f(E)        forward differences; every d/dz term
                                           is  dz/kappa_H(k) + psi_H  with
                                           psi_H = b_H psi_H + c_H dz  (column 1
                                           of the pml arrays, z = k + 1/2)
  E half step  E = ca*E + cb*(curl_b(H) + J(n))   backward differences, d/dz
                                           terms with column 0 (z = k)

  cb = (dt/eps) / (1 + q),  ca = (1 - q)/(1 + q),  q = sigma_abs*dt/(2 eps),
  sigma_abs = absorption_mask[c, i, j];  b = exp(-(sigma/kappa + alpha) dt),
  c = sigma (b - 1) / (sigma kappa + kappa^2 alpha)  (0 where sigma = 0 or
  kappa = inf);  psi exists only for k < pml_lo or k >= zz - pml_hi.

  Fields outside the arrays are zero.  eps outside the ``offset`` window is the
  nearest edge value.  J(n) is source_field times source_waveform[n]:
    x/y plane at p (p even): column 0 drives plane p, column 1 plane p - 1
      (this reading makes pjz's odd-position shift at _field.py:230-233 land on
      the requested plane); transverse components (Ey,Ez) / (Ex,Ez).
    z plane at p: source_field (2 quadratures, 2 comps (Ex,Ey), XX, YY, 1),
      J = sf[0] * w[n,0] + sf[1] * w[n,1].
  Output s in range(*output_steps) is E after the E half step of step s,
  over the eps window at ``offset``: out (nn, 3, xx, yy, ze).

  The source enters with a plus sign (the field of a physical current -J).
  pjz's scatter() gradient, sum Re(g f_i f_j / a_i) (_field.py:380-382,
  393-398), is odd in the sign of the fields, while the S-values are not.
  Under this update the reciprocity term is K dS_ij/deps with K close to
  Omega ww eps_mode exp(-i omega dt / 2) > 0 (Omega = 2 sin(omega dt/2)/dt,
  ww frequencies, eps_mode a modal average of eps), measured against finite
  differences in tests/test_oracle_scatter.py; with -J it would be -K, a
  gradient that points uphill.  pjz's optimisation loop
  (examples/_utils.py:27-29) descends along that gradient, so the sign that
  makes it a descent direction is taken.

Pinning: brute-force per-cell loops (``brute.py``), discrete energy
conservation, closed-form cavity-mode evolution, first-step source closed
form, lossy-mode decay rate, CPML reflection bound, kappa=inf == independent
2-D slices, edge-replicate fill == explicit padding.  These live in
``tests/test_oracle_*.py``.  Parity with the real fdtdz kernel is unpinned:
its source is not available.

Works on numpy arrays, or on torch float64 CPU tensors (``xp=torch``) so that
torch.autograd yields the exact gradient of this discrete map.
"""

import math

import numpy as np


# ---------------------------------------------------------------- utilities
def _cat(xp, arrs, axis):
  if xp is np:
    return np.concatenate(arrs, axis=axis)
  return xp.cat(arrs, dim=axis)


def _take(a, axis, sl):
  idx = [slice(None)] * a.ndim
  idx[axis] = sl
  return a[tuple(idx)]


def dfwd(xp, f, axis):
  """f[i+1] - f[i], with f beyond the last index taken as zero."""
  return _cat(xp, [_take(f, axis, slice(1, None)) - _take(f, axis, slice(0, -1)),
                   -_take(f, axis, slice(-1, None))], axis)


def dbwd(xp, f, axis):
  """f[i] - f[i-1], with f before index 0 taken as zero."""
  return _cat(xp, [_take(f, axis, slice(0, 1)),
                   _take(f, axis, slice(1, None)) - _take(f, axis, slice(0, -1))],
              axis)


def pml_coefficients(kappa, sigma, alpha, dt):
  """Per-z CPML coefficients from (zz, 2) kappa/sigma/alpha arrays.

  Returns dict of (zz,) float64 arrays: ik (1/kappa), b, c for column 0 (E, at
  z = k) and column 1 (H, at z = k + 1/2).
  """
  kappa = np.asarray(kappa, np.float64)
  sigma = np.asarray(sigma, np.float64)
  alpha = np.asarray(alpha, np.float64)
  out = {}
  for col, tag in ((0, "e"), (1, "h")):
    k, s, a = kappa[:, col], sigma[:, col], alpha[:, col]
    ik = np.where(np.isinf(k), 0.0, 1.0 / np.where(np.isinf(k), 1.0, k))
    b = np.exp(-(s * ik + a) * dt)
    den = s * ik + a
    with np.errstate(divide="ignore", invalid="ignore"):
      c = np.where((s == 0) | (ik == 0) | (den == 0), 0.0,
                   s * ik * ik * (b - 1.0) / np.where(den == 0, 1.0, den))
    out["ik" + tag], out["b" + tag], out["c" + tag] = ik, b, c
  return out


def full_epsilon(xp, eps, domain, offset):
  """Edge-replicated (3, XX, YY, ZZ) permittivity from the eps window."""
  idx = []
  for d in range(3):
    n, o, e = domain[d], offset[d], eps.shape[d + 1]
    i = np.clip(np.arange(n) - o, 0, e - 1)
    idx.append(i if xp is np else xp.as_tensor(i))
  return eps[:, idx[0]][:, :, idx[1]][:, :, :, idx[2]]


def source_axis(source_field, domain):
  """0/1/2 for an x/y/z source plane, from the source_field shape."""
  shp = tuple(source_field.shape)
  if len(shp) == 5:
    return 2
  if len(shp) != 4 or shp[0] != 2:
    raise ValueError(f"source_field must be (2,...) 4-D or (2,2,...) 5-D, got {shp}")
  last = shp[1:]
  cand = [a for a in range(3) if last[a] == 1 and domain[a] != 1]
  if len(cand) != 1:
    cand = [a for a in range(3) if last[a] == 1]
  if len(cand) != 1:
    raise ValueError(f"cannot tell the source plane axis from shape {shp}")
  return cand[0]


TRANSVERSE = {0: (1, 2), 1: (0, 2), 2: (0, 1)}


def source_terms(xp, source_field, source_waveform, source_position, domain):
  """List of (component, axis, index, plane, (w0, w1)) describing J.

  ``plane`` is broadcastable against a slab of the domain of width 1 along
  ``axis``; J at step n on that slab is plane * (w0*wave[n,0] + w1*wave[n,1])
  where (w0, w1) select waveform columns.
  """
  axis = source_axis(source_field, domain)
  p = int(source_position)
  terms = []
  comps = TRANSVERSE[axis]
  if axis == 2:
    sf = source_field if source_field.ndim == 5 else source_field[None]
    for q in range(sf.shape[0]):
      for ci, comp in enumerate(comps):
        terms.append((comp, 2, p, sf[q, ci], (1.0, 0.0) if q == 0 else (0.0, 1.0)))
  else:
    for ci, comp in enumerate(comps):
      terms.append((comp, axis, p, source_field[ci], (1.0, 0.0)))
      if p - 1 >= 0:
        terms.append((comp, axis, p - 1, source_field[ci], (0.0, 1.0)))
  return terms


class Simulation:
  """Holds the FP64 state and advances it one step at a time."""

  def __init__(self, epsilon, dt, source_field, source_waveform, source_position,
               absorption_mask, pml_kappa, pml_sigma, pml_alpha, pml_widths,
               offset=(0, 0, 0), xp=np, E0=None, H0=None):
    self.xp = xp
    f64 = (lambda a: np.asarray(a, np.float64)) if xp is np else (
        lambda a: a if isinstance(a, xp.Tensor) else xp.as_tensor(np.asarray(a, np.float64)))
    self.dt = float(dt)
    mask = f64(absorption_mask)
    XX, YY = mask.shape[1], mask.shape[2]
    ZZ = np.asarray(pml_sigma).shape[0]
    self.domain = (XX, YY, ZZ)
    self.lo, self.hi = int(pml_widths[0]), int(pml_widths[1])
    eps = f64(epsilon)
    self.eps_full = full_epsilon(xp, eps, self.domain, offset)
    q = mask[:, :, :, None] * self.dt / (2.0 * self.eps_full)
    self.cb = (self.dt / self.eps_full) / (1.0 + q)
    self.ca = (1.0 - q) / (1.0 + q)
    pc = pml_coefficients(pml_kappa, pml_sigma, pml_alpha, self.dt)
    pz = np.r_[np.arange(self.lo), np.arange(ZZ - self.hi, ZZ)]
    self.pml_z = pz
    cv = (lambda a: a) if xp is np else (lambda a: xp.as_tensor(a))
    self.ikE, self.ikH = cv(pc["ike"]), cv(pc["ikh"])
    self.bE, self.cE = cv(pc["be"][pz]), cv(pc["ce"][pz])
    self.bH, self.cH = cv(pc["bh"][pz]), cv(pc["ch"][pz])
    self.wave = f64(source_waveform)
    self.terms = source_terms(xp, f64(source_field), self.wave, source_position,
                              self.domain)
    zeros = (lambda s: np.zeros(s)) if xp is np else (
        lambda s: xp.zeros(s, dtype=xp.float64))
    self.E = list(E0) if E0 is not None else [zeros(self.domain) for _ in range(3)]
    self.H = list(H0) if H0 is not None else [zeros(self.domain) for _ in range(3)]
    np_ = len(pz)
    self.psiE = [zeros((XX, YY, np_)) for _ in range(2)]
    self.psiH = [zeros((XX, YY, np_)) for _ in range(2)]
    self.n = 0

  # psi lives on the z indices pml_z; scatter it onto a full-z array term.
  def _with_psi(self, t, psi):
    xp, lo, hi, ZZ = self.xp, self.lo, self.hi, self.domain[2]
    parts = []
    if lo:
      parts.append(t[:, :, :lo] + psi[:, :, :lo])
    parts.append(t[:, :, lo:ZZ - hi])
    if hi:
      parts.append(t[:, :, ZZ - hi:] + psi[:, :, lo:])
    return _cat(xp, parts, 2)

  def _pml_part(self, d):
    xp, lo, hi, ZZ = self.xp, self.lo, self.hi, self.domain[2]
    return _cat(xp, [d[:, :, :lo], d[:, :, ZZ - hi:]], 2)

  def _j(self, comp, n):
    """List of (axis, index, slab) source contributions to component comp."""
    out = []
    for (c, axis, idx, plane, (w0, w1)) in self.terms:
      if c != comp:
        continue
      amp = w0 * self.wave[n, 0] + w1 * self.wave[n, 1]
      out.append((axis, idx, plane * amp))
    return out

  def _add_source(self, f, comp, n):
    xp = self.xp
    for axis, idx, slab in self._j(comp, n):
      shp = list(f.shape)
      shp[axis] = 1
      slab = slab.reshape(shp) if hasattr(slab, "reshape") else slab
      f = _cat(xp, [_take(f, axis, slice(0, idx)),
                    _take(f, axis, slice(idx, idx + 1)) + slab,
                    _take(f, axis, slice(idx + 1, None))], axis)
    return f

  def step(self):
    xp, dt, n = self.xp, self.dt, self.n
    Ex, Ey, Ez = self.E
    Hx, Hy, Hz = self.H
    ikH = self.ikH[None, None, :]
    ikE = self.ikE[None, None, :]
    # ---- H half step
    dzEy = dfwd(xp, Ey, 2)
    dzEx = dfwd(xp, Ex, 2)
    if self.lo + self.hi:
      self.psiH[0] = self.bH * self.psiH[0] + self.cH * self._pml_part(dzEy)
      self.psiH[1] = self.bH * self.psiH[1] + self.cH * self._pml_part(dzEx)
      tzEy = self._with_psi(ikH * dzEy, self.psiH[0])
      tzEx = self._with_psi(ikH * dzEx, self.psiH[1])
    else:
      tzEy, tzEx = ikH * dzEy, ikH * dzEx
    Hx = Hx - dt * (dfwd(xp, Ez, 1) - tzEy)
    Hy = Hy - dt * (tzEx - dfwd(xp, Ez, 0))
    Hz = Hz - dt * (dfwd(xp, Ey, 0) - dfwd(xp, Ex, 1))
    # ---- E half step
    dzHy = dbwd(xp, Hy, 2)
    dzHx = dbwd(xp, Hx, 2)
    if self.lo + self.hi:
      self.psiE[0] = self.bE * self.psiE[0] + self.cE * self._pml_part(dzHy)
      self.psiE[1] = self.bE * self.psiE[1] + self.cE * self._pml_part(dzHx)
      tzHy = self._with_psi(ikE * dzHy, self.psiE[0])
      tzHx = self._with_psi(ikE * dzHx, self.psiE[1])
    else:
      tzHy, tzHx = ikE * dzHy, ikE * dzHx
    cx = self._add_source(dbwd(xp, Hz, 1) - tzHy, 0, n)
    cy = self._add_source(tzHx - dbwd(xp, Hz, 0), 1, n)
    cz = self._add_source(dbwd(xp, Hy, 0) - dbwd(xp, Hx, 1), 2, n)
    Ex = self.ca[0] * Ex + self.cb[0] * cx
    Ey = self.ca[1] * Ey + self.cb[1] * cy
    Ez = self.ca[2] * Ez + self.cb[2] * cz
    self.E = [Ex, Ey, Ez]
    self.H = [Hx, Hy, Hz]
    self.n += 1

  def window(self, offset, shape):
    x0, y0, z0 = offset
    xx, yy, ze = shape
    xp = self.xp
    sl = (slice(x0, x0 + xx), slice(y0, y0 + yy), slice(z0, z0 + ze))
    arrs = [e[sl] for e in self.E]
    return np.stack(arrs) if xp is np else xp.stack(arrs)


def output_step_list(output_steps):
  start, stop, step = (int(v) for v in output_steps)
  return list(range(start, stop, step))


def fdtdz(epsilon, dt, source_field, source_waveform, source_position,
          absorption_mask, pml_kappa, pml_sigma, pml_alpha, pml_widths,
          output_steps, use_reduced_precision=False, launch_params=None,
          offset=(0, 0, 0), xp=np):
  """Oracle for the full fdtdz() call: (nn, 3, xx, yy, ze) float64 snapshots."""
  sim = Simulation(epsilon, dt, source_field, source_waveform, source_position,
                   absorption_mask, pml_kappa, pml_sigma, pml_alpha, pml_widths,
                   offset=offset, xp=xp)
  steps = output_step_list(output_steps)
  shape = tuple(epsilon.shape[1:])
  outs = []
  if steps:
    want = set(steps)
    for n in range(steps[-1] + 1):
      sim.step()
      if n in want:
        outs.append(sim.window(offset, shape))
  if xp is np:
    return np.stack(outs) if outs else np.zeros((0, 3) + shape)
  return xp.stack(outs) if outs else xp.zeros((0, 3) + shape, dtype=xp.float64)


def energy(sim, E_prev):
  """Discrete invariant  |H^(n-1/2)|^2 + <eps E^(n-1), E^n>.

  Conserved exactly by the leapfrog when there is no source, no absorber and
  no CPML (kappa uniform, sigma = 0), because curl_b is the transpose of
  curl_f under the zero-outside boundary.  ``sim`` holds E^n, H^(n-1/2).
  """
  h = sum(float((sim.H[c] ** 2).sum()) for c in range(3))
  e = sum(float((sim.eps_full[c] * E_prev[c] * sim.E[c]).sum()) for c in range(3))
  return h + e
