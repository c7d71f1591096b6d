"""Synthetic initial states of the measured configurations, built on the device
from the operators of this package (no reference counterpart: the reference's
benchmarks use uniform or random states, benchmark_godunov.cpp:20-35).

  * random_state          per-operator kernels: rho ~ U[0.5,2], u_k ~ N(0,0.3),
                          p ~ U[0.5,2], counter-based so independent of launch shape;
  * diagonal_implosion    x0+x1+x2 > 0.45 -> (rho, p) = (1, 1) else (0.125, 0.14), u = 0;
  * adjoint_state         the diagonal implosion with the initial density perturbed
                          by N(0, 0.01) noise times a smooth Gaussian bump;
  * density_direction     seeded direction of a finite-difference check: N(0, 1)
                          noise times the same bump, in the density field only.
"""
from __future__ import annotations

from typing import Tuple

import torch

from . import operators as ops

BUMP_CENTRE = (0.2, 0.2, 0.2)
BUMP_SIGMA = 0.1


def _array(ex: ops.ExecSpace, n: int, dtype, index_bits: int) -> Tuple[ops.EulerPrimArrays, int]:
    na = n + 2
    prim = ops.to_mdspan(ops.create_prim_arrays_1d(ex, na**3, dtype), na, na, na, index_bits=index_bits)
    return prim, na


def random_state(ex: ops.ExecSpace, n: int, dtype=torch.float64, seed: int = 1234,
                 index_bits: int = 64) -> ops.EulerPrimArrays:
    prim, _ = _array(ex, n, dtype, index_bits)
    ops.init_random(ex, prim, seed=seed)
    return prim


def diagonal_implosion(ex: ops.ExecSpace, n: int, dtype=torch.float64, index_bits: int = 64,
                       offset: float = 0.45) -> ops.EulerPrimArrays:
    prim, _ = _array(ex, n, dtype, index_bits)
    ops.init_diagonal(ex, prim, ops.UniformMesh3d(1.0 / n, 1.0 / n, 1.0 / n), offset=offset)
    return prim


def gaussian_bump(n: int, device, dtype=torch.float64) -> torch.Tensor:
    """exp(-|x - c|^2 / (2 sigma^2)) at the cell centres ((i-1)+0.5)/n of the
    (n+2)^3 array, ghosts included; shape (n2, n1, n0)."""
    x = (torch.arange(n + 2, dtype=torch.float64, device=device) - 0.5) / n
    cz, cy, cx = BUMP_CENTRE[2], BUMP_CENTRE[1], BUMP_CENTRE[0]
    r2 = ((x[:, None, None] - cz) ** 2 + (x[None, :, None] - cy) ** 2 + (x[None, None, :] - cx) ** 2)
    return torch.exp(-r2 / (2 * BUMP_SIGMA**2)).to(dtype)


def _noise(n: int, seed: int, sigma: float) -> torch.Tensor:
    g = torch.Generator().manual_seed(seed)
    return sigma * torch.randn((n + 2,) * 3, generator=g, dtype=torch.float64)


def adjoint_state(ex: ops.ExecSpace, n: int, seed: int = 1234, index_bits: int = 64) -> ops.EulerPrimArrays:
    """FP64 diagonal implosion with rho += N(0, 0.01) * bump (noise drawn on the
    host from `seed`, cell by cell in array order)."""
    prim = diagonal_implosion(ex, n, torch.float64, index_bits)
    bump = gaussian_bump(n, prim.d.device)
    rho = prim.d.view(n + 2, n + 2, n + 2)
    rho += _noise(n, seed, 0.01).to(prim.d.device) * bump
    return prim


def density_direction(n: int, j: int, device) -> torch.Tensor:
    """Direction j of the finite-difference check (seed 5678 + j): N(0, 1) * bump
    in the density, shape (n2, n1, n0). Velocities stay exactly zero under such
    perturbations, so the u* = 0 and u_L = u_R ties of the resting regions keep
    their branch and the check stays away from switches there."""
    return _noise(n, 5678 + j, 1.0).to(device) * gaussian_bump(n, device)
