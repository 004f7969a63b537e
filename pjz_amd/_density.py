"""Per-pixel density from a raw design variable: pjz's ``density()``.

Mirrors ``src/pjz/_density.py``, the three-field scheme of Zhou et al.
(2015): a conical blur of the variable (``_filter``, ``:17-20``, edge-padded,
kernel ``max(0, r - distance)`` normalised, ``:9-14``), a projection that
binarises every pixel except those on the ``eta`` boundary (``_project``,
``:23-31``), and a minimum-feature-size loss weighted by an inflection
indicator ``exp(-c |grad|^2)`` (``_inflection`` / ``_geom_loss``,
``:34-45``).  Differentiable.  A float32 variable on the GPU runs the HIP
kernels of ``csrc/kernels_density.hip`` (forward and exact reverse pass); a
CPU variable runs the torch restatement below on the host.
"""

import math

import torch
import torch.nn.functional as F


def cone(radius, dtype=torch.float32, device=None):
  """Normalised conical kernel of side 2 ceil(radius - 1/2) + 1 (``:9-14``)."""
  r = math.ceil(radius - 0.5)
  u = (torch.arange(2 * r + 1, dtype=torch.float64) - r) ** 2
  weights = torch.clamp(radius - torch.sqrt(u[:, None] + u[None, :]), min=0)
  return (weights / weights.sum()).to(dtype=dtype, device=device)


def blur(u, radius):
  """Edge-padded cone blur, same shape as ``u`` (``:17-20``)."""
  k = cone(radius, u.dtype, u.device)
  r = k.shape[0] // 2
  padded = F.pad(u[None, None], (r, r, r, r), mode="replicate")
  # The kernel is symmetric, so correlation (conv2d) equals convolution.
  return F.conv2d(padded, k[None, None])[0, 0]


def _edge_pad1(x):
  x = torch.cat([x[:1], x, x[-1:]], dim=0)
  return torch.cat([x[:, :1], x, x[:, -1:]], dim=1)


def boundary_cells(v):
  """Pixels whose sign of ``v >= 0`` differs from a 4-neighbour (edges
  replicated); ``_density.py:23-27``."""
  b = _edge_pad1(v >= 0)
  c = b[1:-1, 1:-1]
  return ((c ^ b[:-2, 1:-1]) | (c ^ b[2:, 1:-1]) | (c ^ b[1:-1, :-2]) | (c ^ b[1:-1, 2:]))


def project(u, eta):
  """Binarise at ``eta`` except on boundary pixels (``:30-31``)."""
  return torch.where(boundary_cells(u - eta), u, (torch.sign(u - eta) + 1) / 2)


def inflection(u, c):
  """exp(-c |central difference|^2), edges replicated (``:34-38``)."""
  p = _edge_pad1(u)
  g2 = (p[2:, 1:-1] - p[:-2, 1:-1]) ** 2 + (p[1:-1, 2:] - p[1:-1, :-2]) ** 2
  return torch.exp(-c * g2)


def geom_loss(uproj, ufilt, c, eta_lo, eta_hi):
  """Minimum-feature-size loss per pixel (``:41-45``)."""
  w = inflection(ufilt, c)
  zero = torch.zeros((), dtype=ufilt.dtype, device=ufilt.device)
  return (uproj * w * torch.minimum(zero, ufilt - eta_hi) ** 2 +
          (1 - uproj) * w * torch.minimum(zero, eta_lo - ufilt) ** 2)


class _DensityKernel(torch.autograd.Function):
  """density() through the C ABI (``pjz_fdtd_density_forward`` /
  ``_backward``) for a float32 device variable."""

  @staticmethod
  def forward(ctx, u, radius, alpha, c, eta, eta_lo, eta_hi):
    from . import _lib
    u = u.detach().contiguous()
    X, Y = u.shape
    dens, loss, uf = (torch.empty_like(u) for _ in range(3))
    with torch.cuda.device(u.device):
      stream = torch.cuda.current_stream(u.device).cuda_stream
      _lib.check(_lib.load().pjz_fdtd_density_forward(
          1, X, Y, radius, alpha, c, eta, eta_lo, eta_hi, u.data_ptr(), dens.data_ptr(),
          loss.data_ptr(), uf.data_ptr(), stream), "pjz_fdtd_density_forward")
    ctx.save_for_backward(uf)
    ctx.args = (radius, alpha, c, eta, eta_lo, eta_hi)
    return dens, loss

  @staticmethod
  def backward(ctx, g_dens, g_loss):
    from . import _lib
    import ctypes
    (uf,) = ctx.saved_tensors
    X, Y = uf.shape
    g_dens = torch.zeros_like(uf) if g_dens is None else g_dens.to(uf.dtype).contiguous()
    g_loss = torch.zeros_like(uf) if g_loss is None else g_loss.to(uf.dtype).contiguous()
    radius = ctx.args[0]
    lib = _lib.load()
    nbytes = ctypes.c_size_t(0)
    _lib.check(lib.pjz_fdtd_density_workspace_bytes(1, X, Y, radius, ctypes.byref(nbytes)),
               "pjz_fdtd_density_workspace_bytes")
    # torch's caching allocator hands out 512-byte aligned blocks.
    ws = torch.empty(max(nbytes.value, 1), dtype=torch.uint8, device=uf.device)
    g_u = torch.empty_like(uf)
    with torch.cuda.device(uf.device):
      stream = torch.cuda.current_stream(uf.device).cuda_stream
      _lib.check(lib.pjz_fdtd_density_backward(
          1, X, Y, *ctx.args, uf.data_ptr(), g_dens.data_ptr(), g_loss.data_ptr(),
          g_u.data_ptr(), ws.data_ptr(), nbytes.value, stream), "pjz_fdtd_density_backward")
    return g_u, None, None, None, None, None, None


def density(u, radius, alpha, c=1.0, eta=0.5, eta_lo=0.25, eta_hi=0.75):
  """pjz's ``density()`` (``_density.py:48-79``): ``(density, loss)``, both
  shaped like the (xx, yy) variable ``u``.  A GPU variable runs the HIP
  kernels (float32 only: they compute in float32); a CPU variable runs the
  torch restatement on the host."""
  u = torch.as_tensor(u)
  if not u.is_floating_point():
    u = u.to(torch.float32)
  if u.ndim != 2:
    raise ValueError(f"u must be (xx, yy), got {tuple(u.shape)}")
  if u.is_cuda:
    if u.dtype != torch.float32:
      raise ValueError(f"density() on the GPU computes in float32, got {u.dtype}")
    return _DensityKernel.apply(u, float(radius), float(alpha), float(c), float(eta),
                                float(eta_lo), float(eta_hi))
  ufilt = blur(u, radius)
  uproj = project(ufilt, eta)
  return alpha * uproj + (1 - alpha) * ufilt, geom_loss(uproj, ufilt, c, eta_lo, eta_hi)
