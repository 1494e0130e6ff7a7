"""Shared helpers: small problems and seeded states (tests only)."""
import torch

from difffea_amd.problem import build_problem

DTYPE = torch.float64


def unit_square(element_type="quad4", n=8, mode="displacement", value=(0.1, None), E=1.0, nu=0.3, **kw):
    spec = dict(element_type=element_type, nx=n, ny=kw.pop("ny", n), E=E, nu=nu, load_mode=mode,
                load_value=value, fixed_side="left", load_side="right", **kw)
    return build_problem(spec)


def cantilever(element_type="quad4", h=0.5, load=(0.0, -40.0), E=2600.0, nu=0.3, **kw):
    """The reference's small test problem: 2x1 rectangle, clamped left, consistent force on the right
    (other materials with their library defaults, as the reference's helper builds them)."""
    spec = dict(element_type=element_type, geometry="rectangle", geom_params=dict(width=2.0, height=1.0), h=h,
                E=E, nu=nu, load_mode="force", load_value=load, fixed_side="left", load_side="right", **kw)
    return build_problem(spec)


def state(params, seed=0, scale=0.02):
    """A non-trivial linearisation point and a random direction (same recipe as the reference tests)."""
    g = torch.Generator().manual_seed(seed)
    n = params["f"].shape[0]
    u = scale * torch.randn(n, dtype=DTYPE, generator=g) * params["m"]
    v = torch.randn(n, dtype=DTYPE, generator=g)
    return u, v


def stretch_state(params, h, seed=0, stretch=0.05):
    """u = (F0 - I) X with F0 = diag(1+stretch, 1), plus N(0, (1e-3 h)^2) per dof."""
    g = torch.Generator().manual_seed(seed)
    X = params["nodes"]
    u = torch.zeros_like(X)
    u[:, 0] = stretch * X[:, 0]
    u = u.reshape(-1) + 1e-3 * h * torch.randn(u.numel(), dtype=DTYPE, generator=g)
    return u


def rel(a, b):
    """max|a-b| / max|b| (the reference's agreement metric, test_tangent_operators.py:26-27)."""
    return float((a.double().cpu() - b.double().cpu()).abs().max() / b.double().cpu().abs().max())


def dense(op, n):
    eye = torch.eye(n, dtype=DTYPE)
    return torch.stack([op(eye[i]) for i in range(n)], dim=1)
