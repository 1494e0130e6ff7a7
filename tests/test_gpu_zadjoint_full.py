"""Adjoint sensitivity at the full 256x256 Q4 size (65,536 elements, 132,098 dofs): rho_e ~ U(0.1, 1)
(seed 3), left edge clamped, dead vertical load on the right edge with total 0.05 * mean(rho) (E = 1, H = 1),
compliance f.u, FP64.  The gradient from the hand-written kernel is checked along random directions
dtheta_e = +-1e-3 rho_e (seed 4) against central differences of the compliance: the directional change is
~1e-3 of the compliance, while solves to ||R|| <= 1e-10 ||f|| pin the compliance to ~1e-9 relative, so the
difference quotient resolves the directional derivative to ~1e-5 (bar 1e-4).  The per-element gradient
itself is pinned against the reference's torch.func adjoint at 16x16 (golden vector G3) and against the
oracle in test_gpu_solve.py."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_adjoint_gradient_full_size_matches_directional_differences():
    from difffea_amd.adjoint import compute_sensitivity
    from difffea_amd.newton_krylov import HipNewton
    from difffea_amd.problem import build_problem
    n = 256
    rho = 0.1 + 0.9 * torch.rand(n * n, dtype=torch.float64, generator=torch.Generator().manual_seed(3))
    F = 0.05 * float(rho.mean())
    p = build_problem(dict(element_type="quad4", nx=n, ny=n, E=1.0, nu=0.3, load_mode="force",
                           load_value=(0.0, -F)))
    f = p["f"].cuda()
    tol = 1e-10 * float(torch.norm(p["f"]))
    u, g = compute_sensitivity(p, rho.cuda(), lambda w: torch.dot(f, w), n_load_steps=1, tol=tol,
                               adjoint_rtol=1e-10, cg_max_iter=20000, adjoint_max_iter=20000)
    assert torch.isfinite(g).all() and g.shape == (n * n,)
    print(f"adjoint gradient done, |g| = {float(torch.norm(g)):.6e}", flush=True)
    rng = np.random.default_rng(4)
    for _ in range(2):
        d = torch.from_numpy(rng.choice([-1.0, 1.0], n * n)) * 1e-3 * rho
        L = []
        for sgn in (1.0, -1.0):
            up, info = HipNewton(p, (rho + sgn * d).cuda()).solve(1, tol, cg_max_iter=20000)
            assert info["converged"]
            L.append(float(torch.dot(f, up)))
            print(f"perturbed solve: {info['total_newton']} Newton, {info['total_cg']} CG", flush=True)
        fd = (L[0] - L[1]) / 2.0
        an = float(torch.dot(g.cpu(), d))
        assert abs(an - fd) <= 1e-4 * abs(fd), (an, fd)
