"""Per-cell triple-loop stepper -- TEST INFRASTRUCTURE ONLY.

Restates the update of ``fdtd_spec`` cell by cell from the Yee positions, with
explicit neighbour lookups (zero outside the arrays), so that a transposed
index, a wrong difference direction or a misplaced CPML column in the
vectorised oracle shows up as a mismatch.  Pure Python: only for grids of a
few hundred cells.

Yee positions:  Ex (i+1/2, j, k)   Ey (i, j+1/2, k)   Ez (i, j, k+1/2)
                Hx (i, j+1/2, k+1/2)  Hy (i+1/2, j, k+1/2)  Hz (i+1/2, j+1/2, k)
"""

import math

import numpy as np


def run(eps_full, mask, dt, kappa, sigma, alpha, pml_widths, src_axis, src_pos,
        src_field, wave, nsteps):
  """Advance nsteps from zero fields; returns (E, H) float64 (3, X, Y, Z)."""
  X, Y, Z = eps_full.shape[1:]
  lo, hi = pml_widths
  E = np.zeros((3, X, Y, Z))
  H = np.zeros((3, X, Y, Z))
  psiE = np.zeros((2, X, Y, Z))  # only PML cells ever become nonzero
  psiH = np.zeros((2, X, Y, Z))

  def coef(col, k):
    kap, sig, alp = kappa[k, col], sigma[k, col], alpha[k, col]
    ik = 0.0 if math.isinf(kap) else 1.0 / kap
    b = math.exp(-(sig * ik + alp) * dt)
    den = sig * ik + alp
    c = 0.0 if (sig == 0 or ik == 0 or den == 0) else sig * ik * ik * (b - 1) / den
    return ik, b, c

  in_pml = lambda k: k < lo or k >= Z - hi

  def g(F, c, i, j, k):
    if 0 <= i < X and 0 <= j < Y and 0 <= k < Z:
      return F[c, i, j, k]
    return 0.0

  def J(c, i, j, k, n):
    tot = 0.0
    if src_axis == 2:
      sf = src_field if src_field.ndim == 5 else src_field[None]
      if k == src_pos and c in (0, 1):
        for q in range(sf.shape[0]):
          tot += sf[q, c, i, j, 0] * wave[n, q]
      return tot
    comps = (1, 2) if src_axis == 0 else (0, 2)
    if c not in comps:
      return 0.0
    ci = comps.index(c)
    pos = (i, j, k)[src_axis]
    idx = [i, j, k]
    idx[src_axis] = 0
    val = src_field[ci, idx[0], idx[1], idx[2]]
    if pos == src_pos:
      tot += val * wave[n, 0]
    if pos == src_pos - 1:
      tot += val * wave[n, 1]
    return tot

  for n in range(nsteps):
    Hn = H.copy()
    for i in range(X):
      for j in range(Y):
        for k in range(Z):
          ik, b, c = coef(1, k)
          dzEy = g(E, 1, i, j, k + 1) - g(E, 1, i, j, k)
          dzEx = g(E, 0, i, j, k + 1) - g(E, 0, i, j, k)
          if in_pml(k):
            psiH[0, i, j, k] = b * psiH[0, i, j, k] + c * dzEy
            psiH[1, i, j, k] = b * psiH[1, i, j, k] + c * dzEx
          tEy = ik * dzEy + (psiH[0, i, j, k] if in_pml(k) else 0.0)
          tEx = ik * dzEx + (psiH[1, i, j, k] if in_pml(k) else 0.0)
          dyEz = g(E, 2, i, j + 1, k) - g(E, 2, i, j, k)
          dxEz = g(E, 2, i + 1, j, k) - g(E, 2, i, j, k)
          dxEy = g(E, 1, i + 1, j, k) - g(E, 1, i, j, k)
          dyEx = g(E, 0, i, j + 1, k) - g(E, 0, i, j, k)
          Hn[0, i, j, k] = H[0, i, j, k] - dt * (dyEz - tEy)
          Hn[1, i, j, k] = H[1, i, j, k] - dt * (tEx - dxEz)
          Hn[2, i, j, k] = H[2, i, j, k] - dt * (dxEy - dyEx)
    H = Hn
    En = E.copy()
    for i in range(X):
      for j in range(Y):
        for k in range(Z):
          ik, b, c = coef(0, k)
          dzHy = g(H, 1, i, j, k) - g(H, 1, i, j, k - 1)
          dzHx = g(H, 0, i, j, k) - g(H, 0, i, j, k - 1)
          if in_pml(k):
            psiE[0, i, j, k] = b * psiE[0, i, j, k] + c * dzHy
            psiE[1, i, j, k] = b * psiE[1, i, j, k] + c * dzHx
          tHy = ik * dzHy + (psiE[0, i, j, k] if in_pml(k) else 0.0)
          tHx = ik * dzHx + (psiE[1, i, j, k] if in_pml(k) else 0.0)
          curl = (
              (g(H, 2, i, j, k) - g(H, 2, i, j - 1, k)) - tHy,
              tHx - (g(H, 2, i, j, k) - g(H, 2, i - 1, j, k)),
              (g(H, 1, i, j, k) - g(H, 1, i - 1, j, k))
              - (g(H, 0, i, j, k) - g(H, 0, i, j - 1, k)),
          )
          for cc in range(3):
            e = eps_full[cc, i, j, k]
            q = mask[cc, i, j] * dt / (2 * e)
            ca = (1 - q) / (1 + q)
            cb = (dt / e) / (1 + q)
            En[cc, i, j, k] = ca * E[cc, i, j, k] + cb * (curl[cc] + J(cc, i, j, k, n))
    E = En
  return E, H
