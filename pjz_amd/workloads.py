"""Synthetic fdtdz() inputs for the benchmark and parity workloads.

Each builder returns the keyword arguments of :func:`pjz_amd.fdtdz.fdtdz`
(float32 numpy arrays plus static arguments), generated from fixed seeds.

  layered_64     64x64x64, 2000 steps: SiO2 (eps 2.085) below z=24, a Si
                 (12.25) strip 24<=z<32, 26<=y<38, air above, +-0.01 seeded
                 noise; Ey Gaussian-pulse current plane at x=10 (centre
                 frequency 2pi/20, width 60 steps, centred at step 240);
                 z-CPML 10 cells/side (cubic sigma to 0.5, kappa 1, alpha
                 0.05); absorber 10 cells; snapshots at steps 1990..1999.
  native_512     512x512x128, 10000 steps: the same stack scaled (SiO2 below
                 z=48, core slab 48<=z<64 with a 48-cell strip and a 200x200
                 seeded Bernoulli(0.5) {2.085, 12.25} design region at
                 x,y in [156,356)); seeded Gaussian yz mode-profile source at
                 x=40 driven by pjz's ramped sine at 2pi/20; z-CPML 16 cells
                 (pjz default sigma), absorber 32 cells; 8 snapshots.
"""

import math

import numpy as np

from . import boundaries as B


def _pulse(tt, omega, t0, width):
  t = np.arange(tt, dtype=np.float64)
  return np.sin(omega * 0.5 * (t - t0)) * np.exp(-((t - t0) ** 2) / (2.0 * width ** 2))


def layered_64(steps=2000, seed=0):
  X = Y = Z = 64
  rng = np.random.default_rng(seed)
  eps = np.ones((3, X, Y, Z))
  eps[:, :, :, :24] = 2.085
  eps[:, :, 26:38, 24:32] = 12.25
  eps += rng.uniform(-0.01, 0.01, eps.shape)
  sf = np.zeros((2, 1, Y, Z))
  sf[0] = 1.0  # Ey current over the whole x = 10 plane
  wave = np.zeros((steps, 2))
  wave[:, 0] = _pulse(steps, 2 * math.pi / 20, 240.0, 60.0)
  sig = B.pml_sigma((10, 10), Z, 0.125, 3.0)
  return dict(
      epsilon=eps.astype(np.float32), dt=0.5, source_field=sf.astype(np.float32),
      source_waveform=wave.astype(np.float32), source_position=10,
      absorption_mask=B.absorption_mask(X, Y, 10, 2.5e-3),
      pml_kappa=np.ones((Z, 2), np.float32), pml_sigma=sig,
      pml_alpha=np.full((Z, 2), 0.05, np.float32), pml_widths=(10, 10),
      output_steps=(steps - 10, steps, 3), use_reduced_precision=False,
      launch_params=None, offset=(0, 0, 0))


def native_512_z64(steps=10000, seed=0):
  """native_512 at half height, 512x512x64 (one cell per lane: the height
  where three steps per launch pay)."""
  return native_512(steps=steps, seed=seed, zz=64)


def native_512_z96(steps=10000, seed=0):
  """native_512 at 512x512x96, the height pjz's default SimParams give
  (zz = 128 - 32)."""
  return native_512(steps=steps, seed=seed, zz=96)


def native_512(steps=10000, seed=0, xx=512, yy=512, zz=128, design=200):
  """The 512x512x128 workload described above; ``xx``, ``yy``, ``zz`` and
  the ``design`` square's side shrink it for parity runs the oracle can
  finish (same stack, source, boundaries and snapshot schedule)."""
  X, Y, Z = xx, yy, zz
  rng = np.random.default_rng(seed)
  zs = Z / 128.0
  zc0, zc1 = int(48 * zs), int(64 * zs)
  eps = np.ones((3, X, Y, Z), np.float32)
  eps[:, :, :, :zc0] = 2.085
  yc = Y // 2
  eps[:, :, yc - 24:yc + 24, zc0:zc1] = 12.25
  d0x, d0y = X // 2 - design // 2, Y // 2 - design // 2
  pattern = np.where(rng.random((design, design)) < 0.5, 2.085, 12.25).astype(np.float32)
  eps[:, d0x:d0x + design, d0y:d0y + design, zc0:zc1] = pattern[None, :, :, None]
  yy_, zz_ = np.meshgrid(np.arange(Y) - yc, np.arange(Z) - (zc0 + zc1) / 2, indexing="ij")
  prof = np.exp(-(yy_ ** 2) / (2 * 16.0 ** 2) - (zz_ ** 2) / (2 * 6.0 ** 2))
  prof *= 1 + 0.05 * rng.standard_normal(prof.shape)
  sf = np.zeros((2, 1, Y, Z), np.float32)
  sf[0, 0] = prof
  rs = B.ramped_sin(np.array([2 * math.pi / 20]), 4.0, 4.0, 0.5, steps)
  wave = np.zeros((steps, 2), np.float32)
  wave[:, 0] = np.imag(rs)
  return dict(
      epsilon=eps, dt=0.5, source_field=sf, source_waveform=wave, source_position=40,
      absorption_mask=B.absorption_mask(X, Y, 32, 2.5e-4),
      pml_kappa=np.ones((Z, 2), np.float32), pml_sigma=B.pml_sigma((16, 16), Z, 0.5, 1.3),
      pml_alpha=np.zeros((Z, 2), np.float32), pml_widths=(16, 16),
      output_steps=(steps - 24, steps, 3), use_reduced_precision=True,
      launch_params=None, offset=(0, 0, 0))


def cells(kw):
  m = np.shape(kw["absorption_mask"])
  return m[1] * m[2] * np.shape(kw["pml_sigma"])[0]


def time_steps(kw):
  """Number of FDTD steps fdtdz() runs: the last snapshot step + 1."""
  s = range(*kw["output_steps"])
  return s[-1] + 1 if len(s) else 0


def batch64(steps=4000, seed=0, n=64, device=None, sizes=(128, 192, 256, 384, 512),
            heights=(64, 128)):
  """Uneven batch: n simulations, xx, yy drawn from ``sizes`` and zz from
  ``heights`` (seeded), each with its own Bernoulli(0.5) {2.085, 12.25} design
  region over 25% of the xy area (a centred square) and 8 z layers, a
  Gaussian-profile Ey current plane at x = 16 driven by a Gaussian pulse
  (centre frequency 2pi/20), z-CPML 10 cells per side, absorber 16 cells,
  ``steps`` steps and 4 snapshots.  Arrays are built with torch on ``device``
  (the GPU by default) so the batch does not pass through host memory.
  """
  import torch
  rng = np.random.default_rng(seed)
  dev = device or torch.device("cuda", 0)
  gen = torch.Generator(device=dev).manual_seed(seed)
  calls = []
  for i in range(n):
    X, Y = (int(v) for v in rng.choice(sizes, 2))
    Z = int(rng.choice(heights))
    eps = torch.ones((3, X, Y, Z), dtype=torch.float32, device=dev)
    zc0 = Z // 2 - 4
    eps[:, :, :, : Z // 3] = 2.085
    dx, dy = X // 4, Y // 4  # half side of a square holding 25% of the xy area
    cx, cy = X // 2, Y // 2
    pat = torch.rand((2 * dx, 2 * dy), generator=gen, device=dev) < 0.5
    design = torch.where(pat, 2.085, 12.25).to(torch.float32)
    eps[:, cx - dx:cx + dx, cy - dy:cy + dy, zc0:zc0 + 8] = design[None, :, :, None]
    yy_ = torch.arange(Y, device=dev, dtype=torch.float32)[:, None] - cy
    zz_ = torch.arange(Z, device=dev, dtype=torch.float32)[None, :] - (zc0 + 4)
    sf = torch.zeros((2, 1, Y, Z), dtype=torch.float32, device=dev)
    sf[0, 0] = torch.exp(-(yy_ ** 2) / (2 * 12.0 ** 2) - (zz_ ** 2) / (2 * 6.0 ** 2))
    wave = np.zeros((steps, 2), np.float32)
    wave[:, 0] = _pulse(steps, 2 * math.pi / 20, 200.0, 50.0)
    calls.append(dict(
        epsilon=eps, dt=0.5, source_field=sf, source_waveform=torch.as_tensor(wave, device=dev),
        source_position=16, absorption_mask=torch.as_tensor(B.absorption_mask(X, Y, 16, 1e-3),
                                                            device=dev),
        pml_kappa=torch.ones((Z, 2), device=dev),
        pml_sigma=torch.as_tensor(B.pml_sigma((10, 10), Z, 0.5, 3.0), device=dev),
        pml_alpha=torch.full((Z, 2), 0.05, device=dev), pml_widths=(10, 10),
        output_steps=(steps - 40, steps, 10), use_reduced_precision=False,
        launch_params=None, offset=(0, 0, 0)))
  return calls


GRAD_OMEGAS = (2 * math.pi / 18, 2 * math.pi / 20, 2 * math.pi / 22)


def grad_256(steps=6000, seed=0, xx=256, yy=256, zz=128, monitor_x=220):
  """Gradient workload: the layered stack of ``layered_64`` scaled to
  256x256x128 (SiO2 below z=48, a Si strip 48<=z<64, 104<=y<152), +-0.01
  seeded noise, an Ey current plane at x=40 driven by pjz's ramped sine
  averaged over 2pi/18, 2pi/20, 2pi/22; z-CPML 10 cells, absorber 16 cells;
  snapshots on pjz's sampling schedule for the three frequencies.

  Returns (kwargs, mode, omegas): ``mode`` is a seeded Gaussian (Ey, Ez)
  profile (2, yy, zz) on the plane x = ``monitor_x`` used by
  :func:`overlap_objective`.
  """
  rng = np.random.default_rng(seed)
  X, Y, Z = xx, yy, zz
  s = zz / 64.0
  eps = np.ones((3, X, Y, Z), np.float32)
  eps[:, :, :, : int(24 * s)] = 2.085
  y0, y1 = int(26 * yy / 64), int(38 * yy / 64)
  eps[:, :, y0:y1, int(24 * s):int(32 * s)] = 12.25
  eps += rng.uniform(-0.01, 0.01, eps.shape).astype(np.float32)
  sf = np.zeros((2, 1, Y, Z), np.float32)
  yy_, zz_ = np.meshgrid(np.arange(Y) - (y0 + y1) / 2, np.arange(Z) - 28 * s, indexing="ij")
  sf[0, 0] = np.exp(-(yy_ ** 2) / (2 * (0.5 * (y1 - y0)) ** 2) - (zz_ ** 2) / (2 * (4 * s) ** 2))
  omegas = np.array(GRAD_OMEGAS)
  rs = B.ramped_sin(omegas, 4.0, 4.0, 0.5, steps)
  wave = np.zeros((steps, 2), np.float32)
  wave[:, 0] = np.imag(rs)
  interval = B.sampling_interval(omegas.min(), omegas.max(), 3, 0.5)
  kw = dict(epsilon=eps, dt=0.5, source_field=sf, source_waveform=wave, source_position=40,
            absorption_mask=B.absorption_mask(X, Y, 16, 1e-3),
            pml_kappa=np.ones((Z, 2), np.float32), pml_sigma=B.pml_sigma((10, 10), Z, 0.5, 3.0),
            pml_alpha=np.full((Z, 2), 0.05, np.float32), pml_widths=(10, 10),
            output_steps=(steps - 2 * interval * 3 - 1, steps, interval),
            use_reduced_precision=False, launch_params=None, offset=(0, 0, 0))
  mode = np.zeros((2, Y, Z), np.float32)
  cy, cz = (y0 + y1) / 2 + rng.uniform(-2, 2), 28 * s + rng.uniform(-2, 2)
  mode[0] = np.exp(-((np.arange(Y)[:, None] - cy) ** 2) / (2 * 8.0 ** 2)
                   - ((np.arange(Z)[None, :] - cz) ** 2) / (2 * 5.0 ** 2))
  mode[1] = 0.1 * rng.standard_normal((Y, Z)) * mode[0]
  return kw, mode, omegas


def overlap_objective(fields, mode, omegas, output_steps, dt=0.5, monitor_x=220):
  """sum over omega of |<E(omega) on plane monitor_x, mode>|^2 (torch).

  ``fields`` are fdtdz() snapshots (nn, 3, xx, yy, zz); the projection onto
  frequencies is field.project (pjz's pinv of the output phases).
  """
  import torch
  from ._field import project
  # (ww, 2, yy, zz) complex Ey, Ez on the monitor plane (the projection is
  # linear per cell, so only that plane is projected)
  plane = project(fields[:, 1:, monitor_x], omegas, output_steps, dt)
  m = torch.as_tensor(mode, device=fields.device).to(plane.dtype)
  ov = torch.sum(plane * m[None], dim=(1, 2, 3))
  return torch.sum(torch.abs(ov) ** 2)


def design_batch(steps=2000, seed=0, variants=16, xx=128, yy=128, zz=64, device=None,
                 design=None):
  """Batched inverse-design iteration: ``variants`` design variants x 2 ports
  x 2 frequencies = 4 * variants simulations of one size.

  Each variant's design region (a centred square of side ``design``, by
  default 5/16 of x and y, and 8 z layers in the core; 160 at 384x384x128,
  the full-size iteration) holds eps = 2.085 + (12.25 - 2.085) * sigmoid(latent) with a
  seeded uniform(-3, 3) latent, over a SiO2 (z < zz/3) / air stack with a Si
  strip through the design.  Port 0 injects an Ey Gaussian-profile current
  plane at x = 16, port 1 at x = xx - 18; frequency k drives pjz's ramped sine
  at 2pi/18 or 2pi/22 and samples pjz's 3-snapshot schedule for it.  z-CPML
  10 cells, absorber 12 cells.

  Returns (calls, variant_of, monitors): the fdtdz() keyword dicts (epsilon
  tensors built with torch on ``device``, the GPU by default), the variant of
  each call, and per call (omega, monitor x, mode (2, yy, zz)) for
  :func:`overlap_objective`.
  """
  import torch
  rng = np.random.default_rng(seed)
  dev = device or torch.device("cuda", 0)
  X, Y, Z = xx, yy, zz
  zc0 = Z // 2 - 4
  dx, dy = ((5 * X) // 32, (5 * Y) // 32) if design is None else (design // 2, design // 2)
  cx, cy = X // 2, Y // 2
  base = np.ones((3, X, Y, Z), np.float32)
  base[:, :, :, : Z // 3] = 2.085
  base[:, :, cy - 6:cy + 6, zc0:zc0 + 8] = 12.25
  yy_, zz_ = np.meshgrid(np.arange(Y) - cy, np.arange(Z) - (zc0 + 4), indexing="ij")
  prof = np.exp(-(yy_ ** 2) / (2 * 6.0 ** 2) - (zz_ ** 2) / (2 * 4.0 ** 2)).astype(np.float32)
  sf = np.zeros((2, 1, Y, Z), np.float32)
  sf[0, 0] = prof
  mode = np.zeros((2, Y, Z), np.float32)
  mode[0] = prof
  omegas = (2 * math.pi / 18, 2 * math.pi / 22)
  mask = B.absorption_mask(X, Y, 12, 2e-3)
  sigma = B.pml_sigma((10, 10), Z, 0.5, 3.0)
  calls, variant_of, monitors = [], [], []
  for v in range(variants):
    latent = rng.uniform(-3, 3, (2 * dx, 2 * dy)).astype(np.float32)
    eps = base.copy()
    eps[:, cx - dx:cx + dx, cy - dy:cy + dy, zc0:zc0 + 8] = (
        2.085 + (12.25 - 2.085) / (1 + np.exp(-latent)))[None, :, :, None]
    eps_t = torch.as_tensor(eps, device=dev)
    for port, (src_x, mon_x) in enumerate(((16, X - 24), (X - 18, 22))):
      for w in omegas:
        interval = B.sampling_interval(w, w, 1, 0.5)
        wave = np.zeros((steps, 2), np.float32)
        wave[:, 0] = np.imag(B.ramped_sin(np.array([w]), 4.0, 4.0, 0.5, steps))
        calls.append(dict(
            epsilon=eps_t, dt=0.5, source_field=torch.as_tensor(sf, device=dev),
            source_waveform=torch.as_tensor(wave, device=dev), source_position=src_x,
            absorption_mask=torch.as_tensor(mask, device=dev),
            pml_kappa=torch.ones((Z, 2), device=dev),
            pml_sigma=torch.as_tensor(sigma, device=dev),
            pml_alpha=torch.full((Z, 2), 0.05, device=dev), pml_widths=(10, 10),
            output_steps=(steps - 2 * interval - 1, steps, interval),
            use_reduced_precision=False, launch_params=None, offset=(0, 0, 0)))
        variant_of.append(v)
        monitors.append((w, mon_x, mode))
  return calls, variant_of, monitors
