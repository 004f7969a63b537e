"""Discrete adjoint of the stepper, numpy FP64 -- TEST INFRASTRUCTURE ONLY.

Gradient of L = sum_s <g_s, out_s> with respect to ``epsilon`` for the update
of ``fdtd_spec``.  One forward step is
    psiH' = bH psiH + cH Dzf E            (PML cells)
    H'    = H - dt (Cf E)                 (z terms ikH Dzf E + psiH')
    psiE' = bE psiE + cE Dzb H'
    E'    = ca E + cb K,  K = Cb H' + J   (z terms ikE Dzb H' + psiE')
and the reverse step, with lK = cb lE', runs
    kernel 1 (H cells):  lH_tot = lH' + Cb^T lK + psiE terms;  u = lpsiH' +- dt lH_tot
    kernel 2 (E cells):  lE = ca lE' + dt Cf^T-gather(lH_tot) + psiH terms;
                         lpsiE = bE (lpsiE' -+ lK);  A += lE' E^n;  B += lE' E^(n+1)
where the psi_H cotangent is kept as u and scaled by bH when the next reverse
step reads it.  With gca = A and gcb = (B - ca A)/cb (since cb K = E' - ca E):
    dL/deps = gca dca/deps + gcb dcb/deps,  dca/deps = 2q/(eps (1+q)^2),
    dcb/deps = -dt/(eps^2 (1+q)^2),  summed onto the eps window (edge fill).
Pinning: equals torch.autograd through ``fdtd_spec`` (xp=torch) to 1e-10
relative (tests/test_oracle_adjoint.py), plus a dot-product identity.
"""

import numpy as np

from . import fdtd_spec as S


def _df(f, axis):
  return S.dfwd(np, f, axis)


def _db(f, axis):
  return S.dbwd(np, f, axis)


def grad_epsilon(epsilon, dt, source_field, source_waveform, source_position,
                 absorption_mask, pml_kappa, pml_sigma, pml_alpha, pml_widths,
                 output_steps, g_out, offset=(0, 0, 0), use_reduced_precision=False,
                 launch_params=None, return_out=False):
  eps = np.asarray(epsilon, np.float64)
  sim = S.Simulation(eps, dt, source_field, source_waveform, source_position,
                     absorption_mask, pml_kappa, pml_sigma, pml_alpha, pml_widths,
                     offset=offset)
  steps = S.output_step_list(output_steps)
  nsteps = steps[-1] + 1 if steps else 0
  Es = [np.stack(sim.E)]
  outs = []
  for n in range(nsteps):
    sim.step()
    Es.append(np.stack(sim.E))
    if n in steps:
      outs.append(sim.window(offset, eps.shape[1:]))
  X, Y, Z = sim.domain
  lo, hi = sim.lo, sim.hi
  pz = sim.pml_z
  pc = S.pml_coefficients(pml_kappa, pml_sigma, pml_alpha, dt)
  ikE, bE, cE = pc["ike"], pc["be"], pc["ce"]
  ikH, bH, cH = pc["ikh"], pc["bh"], pc["ch"]
  ca, cb = sim.ca, sim.cb

  def full(p):  # psi-shaped (X, Y, NP) -> (X, Y, Z), zero outside the PML
    f = np.zeros((X, Y, Z))
    f[:, :, pz] = p
    return f

  g_out = np.asarray(g_out, np.float64)
  slot = {s: i for i, s in enumerate(steps)}
  x0, y0, z0 = offset
  ex, ey, ez = eps.shape[1:]
  win = (slice(None), slice(x0, x0 + ex), slice(y0, y0 + ey), slice(z0, z0 + ez))
  lE = np.zeros((3, X, Y, Z))
  lH = np.zeros((3, X, Y, Z))
  u = np.zeros((2, X, Y, len(pz)))    # psi_H cotangent, before the bH scaling
  lpE = np.zeros((2, X, Y, len(pz)))  # psi_E cotangent
  A = np.zeros((3, X, Y, Z))
  Bacc = np.zeros((3, X, Y, Z))
  for n in reversed(range(nsteps)):
    if n in slot:
      lE[win] += g_out[slot[n]]
    # kernel 1 (H cells)
    lK = cb * lE
    pEx = lpE[0] - lK[0][:, :, pz]
    pEy = lpE[1] + lK[1][:, :, pz]
    vX = full(pEx) * cE
    vY = full(pEy) * cE
    lHx = lH[0] - _df(ikE * lK[1], 2) + _df(lK[2], 1) - _df(vY, 2)
    lHy = lH[1] + _df(ikE * lK[0], 2) - _df(lK[2], 0) - _df(vX, 2)
    lHz = lH[2] - _df(lK[0], 1) + _df(lK[1], 0)
    u = np.stack([bH[pz] * u[0] + dt * lHx[:, :, pz],
                  bH[pz] * u[1] - dt * lHy[:, :, pz]])
    # kernel 2 (E cells)
    A += lE * Es[n]
    Bacc += lE * Es[n + 1]
    UX = full(u[0]) * cH
    UY = full(u[1]) * cH
    nEx = ca[0] * lE[0] + dt * (_db(ikH * lHy, 2) - _db(lHz, 1)) - _db(UY, 2)
    nEy = ca[1] * lE[1] + dt * (-_db(ikH * lHx, 2) + _db(lHz, 0)) - _db(UX, 2)
    nEz = ca[2] * lE[2] + dt * (_db(lHx, 1) - _db(lHy, 0))
    lE = np.stack([nEx, nEy, nEz])
    lpE = np.stack([bE[pz] * pEx, bE[pz] * pEy])
    lH = np.stack([lHx, lHy, lHz])
  gca = A
  gcb = (Bacc - ca * A) / cb
  mask = np.asarray(absorption_mask, np.float64)
  ef = sim.eps_full
  q = mask[:, :, :, None] * dt / (2 * ef)
  g_full = gca * 2 * q / (ef * (1 + q) ** 2) - gcb * dt / (ef ** 2 * (1 + q) ** 2)
  grad = np.zeros_like(eps)
  ix = np.clip(np.arange(X) - x0, 0, ex - 1)
  iy = np.clip(np.arange(Y) - y0, 0, ey - 1)
  iz = np.clip(np.arange(Z) - z0, 0, ez - 1)
  np.add.at(grad, (slice(None), ix[:, None, None], iy[None, :, None], iz[None, None, :]),
            g_full)
  if return_out:
    return grad, np.stack(outs)
  return grad
