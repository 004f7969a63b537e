"""Heights 64 < zz <= 96.  pjz's default SimParams give zz = 128 - 32 = 96
(src/pjz/_field.py:56-58), so every default field() call runs this path.
Multiples of 4 with at most 32 PML cells (96 among them) run the LDS-ring
kernel at two cells per lane; other multiples of 3 (69, 90, 93) run three
cells per lane with two rows per wave.

Checked against the FP64 oracle (relative L2 <= 1e-4) on random problems
with ragged tiles (rows not a multiple of 13, fewer rows than a tile), every
source axis, odd heights and PML widths, z as batch; against the two-cells-
per-lane kernel on the same inputs (same arithmetic per cell); through
field() with pjz's default SimParams; and for the gradient, whose segment
recomputes run the same kernel.
"""

import numpy as np
import pytest
import torch

import test_fdtdz_gpu as T
from oracle import adjoint
from oracle import c_oracle as C
from oracle import field_ref
from pjz_amd import _field as Fd
from pjz_amd import fdtdz_torch as F

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("X,Y,Z,widths", [(24, 20, 96, (6, 5)), (19, 31, 96, (16, 16)),
                                          (24, 20, 90, (6, 5)), (19, 31, 90, (16, 16)),
                                          (17, 9, 93, (4, 7)), (30, 14, 69, (0, 3))])
def test_three_cells_per_lane_matches_oracle(axis, X, Y, Z, widths):
  print(f"case axis={axis} shape={X}x{Y}x{Z} pml={widths}", flush=True)
  kw = T.random_problem(axis, X=X, Y=Y, Z=Z, tt=90, seed=X + Y + axis, widths=widths)
  ref = C.fdtdz(**kw)
  got = T.run_gpu(kw, "stream2")
  assert T.rel_l2(got, ref) <= T.TOL
  if Z % 2 == 0:  # the two-cells-per-lane kernel on the same inputs
    kw2 = dict(kw, launch_params=dict(kernel="stream2", tile_y=8))
    two = F.fdtdz(**kw2).cpu().numpy()
    assert T.rel_l2(got, two) <= 1e-6


@pytest.mark.parametrize("Z", [96, 90])
def test_three_cells_per_lane_z_as_batch(Z):
  kw = T.random_problem(0, X=22, Y=27, Z=Z, tt=70, seed=5, widths=(16, 16))
  kw["pml_kappa"] = np.full((Z, 2), np.inf, np.float32)
  kw["pml_sigma"] = np.zeros((Z, 2), np.float32)
  kw["pml_alpha"] = np.zeros((Z, 2), np.float32)
  assert T.rel_l2(T.run_gpu(kw, "stream2"), C.fdtdz(**kw)) <= T.TOL


def test_field_with_default_simparams(gpu):
  """field() with pjz's defaults (50-cell absorber, (16, 16) PML, reduced
  precision: zz = 96) against the oracle's field()."""
  rng = np.random.default_rng(11)
  xx, yy, zz = 20, 16, 24
  eps = (1 + 3 * rng.random((3, xx, yy, zz))).astype(np.float32)
  src = rng.standard_normal((2, 1, yy, zz)).astype(np.float32)
  p = Fd.SimParams(omega_range=(2 * np.pi / 22, 2 * np.pi / 18), tt=500)
  assert Fd.B.zz(p.pml_widths, p.use_reduced_precision) == 96
  omega = np.array([2 * np.pi / 20])
  got = Fd.field(torch.as_tensor(eps, device=gpu), torch.as_tensor(src, device=gpu), omega, 7,
                 p).cpu().numpy()
  ref = field_ref.field(eps, src, omega, 7, p)
  err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
  print(f"field() default SimParams: rel L2 {err:.2e}", flush=True)
  assert err <= 1e-4


@pytest.mark.parametrize("Z", [96, 90])
def test_gradient_at_height_96(Z):
  """The gradient at zz = 96 (the LDS-ring kernels forward and reverse) and
  90 (checkpoint recomputes on the three-cells-per-lane kernel, capturing E
  changes; the reverse steps at two cells per lane)."""
  kw = T.random_problem(0, X=20, Y=17, Z=Z, tt=70, seed=21, widths=(6, 5))
  eps = kw.pop("epsilon")
  rng = np.random.default_rng(3)
  nn = len(range(*kw["output_steps"]))
  g = rng.standard_normal((nn, 3) + eps.shape[1:]).astype(np.float32)
  ref = adjoint.grad_epsilon(eps, g_out=g, **kw)
  e = torch.tensor(eps, device="cuda", requires_grad=True)
  out = F.fdtdz(e, **dict(kw, launch_params=dict(kernel="stream2", ckpt_every=16)))
  (out * torch.as_tensor(g, device="cuda")).sum().backward()
  got = e.grad.cpu().numpy()
  err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
  print(f"gradient zz={Z}: rel L2 {err:.2e}", flush=True)
  assert err <= 1e-3
