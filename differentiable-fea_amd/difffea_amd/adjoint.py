"""Adjoint design sensitivities on the device (`src/difffea/adjoint.py:31-106`, `sensitivity.py:7-38`).

For R(u*, theta) = 0 and a scalar L(u*):   A^T lam = m * dL/du,   dL/dtheta = -lam^T dR/dtheta.

The tangent is symmetric, so the adjoint solve reuses the fused PCG of the forward tangent at u* (in mixed
precision with iterative refinement against an FP64 tangent, so it meets the same tolerance as in FP64).  For the
neo-Hookean material with theta scaling both Lame parameters, R^e_int is linear in theta_e and

    dL/dtheta_e = -sum_{a,i} (m lam)[2 conn(e,a)+i] R^e_int[a,i](u*; mu0, lam0)

which ``dfea_adjoint_grad`` evaluates per element (no scatter, original element order).

``EquilibriumSolve`` is a ``torch.autograd.Function``: forward = the Newton-Krylov solve, backward = the
adjoint solve + the hand-written gradient kernel, so ``loss(u(theta)).backward()`` gives dL/dtheta.
"""
import torch
from torch.func import grad

from .device import HipSystem
from .newton_krylov import HipNewton, MIXED_RTOL_FLOOR


MAX_REFINE = 12       # refinement steps of the mixed-precision adjoint, at most


def solve_adjoint(solver, u_star, dL_du, rtol=1e-6, max_iter=500):
    """PCG on A lam = m * dL/du at u* (A symmetric = A^T); returns (lam, iterations).

    FP64: one PCG to ``rtol``.  Mixed precision: iterative refinement to the same FP64 relative residual --
    the correction solves A d = r run in FP32 (moduli, product, vectors; FP64 dot products) to MIXED_RTOL_FLOOR,
    lam accumulates in FP64, and r = b - A lam comes from an FP64 tangent at u* (its own moduli), so the
    adjoint meets ``rtol`` as the FP64 solve does."""
    s = solver.sys
    s.moduli(u_star)
    m = solver.params["m"].to(u_star.device, torch.float64)
    b = (m * dL_du.to(torch.float64)).contiguous()
    if s.precision != "mixed":
        s.vec["b"].copy_(b)
        s.pcg_start(rtol=rtol, max_iter=max_iter, neg_res=False)
        its, _ = s.pcg_run(max_iter)
        return s.vec["x"].to(torch.float64).clone(), its
    s64 = HipSystem(s.mesh, solver.theta, s.mat, s.f, "fp64", preconditioner=s.preconditioner)
    s64.moduli(u_star)
    return refine(s, s64, b, rtol, max_iter)


def refine(s, s64, b, rtol, max_iter):
    """Mixed-precision iterative refinement of A lam = b: FP32 PCG corrections on s (moduli already set),
    residuals b - A lam with the FP64 tangent s64; returns (lam FP64, total inner iterations)."""
    nb = float(torch.norm(b))
    lam = torch.zeros_like(b)
    r = b
    its = 0
    for _ in range(MAX_REFINE):
        s.vec["b"].copy_(r.to(s.vdt))
        s.pcg_start(rtol=max(rtol, MIXED_RTOL_FLOOR), max_iter=max_iter, neg_res=False)
        k, _ = s.pcg_run(max_iter)
        its += k
        lam += s.vec["x"].to(torch.float64)
        r = b - s64.apply(lam)
        if float(torch.norm(r)) <= rtol * nb:
            break
    return lam, its


def compute_implicit_vjp(solver, u_star, lam):
    """dL/dtheta = -lam^T dR/dtheta at load factor 1 (hand-written kernel)."""
    return solver.sys.adjoint_grad(u_star, lam)


class EquilibriumSolve(torch.autograd.Function):
    """u* = solve(theta); backward: adjoint solve + ``dfea_adjoint_grad``."""

    @staticmethod
    def forward(ctx, theta, params, opts):
        solver = HipNewton(params, theta.detach(), opts.get("precision", "fp64"), theta.device)
        u, info = solver.solve(opts.get("n_load_steps", 5), opts.get("tol", 1e-8), opts.get("max_newton", 25),
                               cg_rtol=opts.get("cg_rtol"), cg_max_iter=opts.get("cg_max_iter", 500))
        if not info["converged"]:
            raise RuntimeError("the equilibrium solve did not converge; no sensitivity is defined")
        ctx.solver, ctx.opts = solver, opts
        ctx.save_for_backward(u)
        ctx.info = info
        return u

    @staticmethod
    def backward(ctx, grad_u):
        (u,) = ctx.saved_tensors
        lam, _ = solve_adjoint(ctx.solver, u, grad_u.to(torch.float64), rtol=ctx.opts.get("adjoint_rtol", 1e-10),
                               max_iter=ctx.opts.get("adjoint_max_iter", 5000))
        return compute_implicit_vjp(ctx.solver, u, lam), None, None


def equilibrium(theta, params, **opts):
    """Differentiable u*(theta) (see ``EquilibriumSolve``)."""
    return EquilibriumSolve.apply(theta, params, opts)


def compute_sensitivity(params, theta, loss_fn, n_load_steps=5, tol=1e-8, max_newton=25, adjoint_rtol=1e-10,
                        precision="fp64", adjoint_max_iter=5000, cg_max_iter=500):
    """(u*, dL/dtheta) like the reference's ``compute_sensitivity`` (solve, grad of the loss, adjoint, vjp)."""
    theta = theta.detach()
    solver = HipNewton(params, theta, precision, theta.device)
    u, info = solver.solve(n_load_steps, tol, max_newton, cg_max_iter=cg_max_iter)
    dL_du = grad(loss_fn)(u)
    lam, _ = solve_adjoint(solver, u, dL_du, rtol=adjoint_rtol, max_iter=adjoint_max_iter)
    return u, compute_implicit_vjp(solver, u, lam)
