"""The fused PCG on a solve large enough that its chunk partials are summed by a separate per-solve kernel
between iterations (more than 2,048 chunks: 800x800 quad4, 2,500 chunks) -- no oracle at this size: the
solution satisfies A x = b to the requested tolerance (checked with the product), the iteration count is
repeatable, it agrees with the textbook PCG run through the product closure, and a served solve of that size
batched with small ones gives the same bits as the single-solve API (the reduction only sums a solve's own
partials)."""
import pytest
import torch

from helpers import rel, stretch_state, unit_square

pytestmark = pytest.mark.gpu
DEV = "cuda"


def test_presummed_pcg_solves_and_is_repeatable():
    from difffea_amd import _lib
    from difffea_amd.newton_krylov import pcg
    from difffea_amd.operators import make_system
    n = 800
    p = unit_square("quad4", n)
    u = stretch_state(p, 1.0 / n, seed=3)
    th = torch.ones(p["conn"].shape[0], dtype=torch.float64)
    s = make_system(u.to(DEV), th.to(DEV), p, "fp64")
    assert s.mesh.plan.n_chunks > 2048
    s.moduli(u.to(DEV))
    b = (torch.randn(u.numel(), dtype=torch.float64, generator=torch.Generator().manual_seed(5)) * p["m"]).to(DEV)
    x1, k1, st1 = s.pcg_solve(b, rtol=1e-3, max_iter=4000)
    x1 = x1.clone()
    x2, k2, st2 = s.pcg_solve(b, rtol=1e-3, max_iter=4000)
    assert st1 == st2 == _lib.PCG_CONVERGED and k1 == k2 and torch.equal(x1, x2)
    assert float(torch.norm(b - s.apply(x1))) <= 1e-3 * float(torch.norm(b)) * 1.0001
    # the closure PCG (textbook recurrences, torch dot products) through the same product and preconditioner
    xr, kr = pcg(s.apply, b, s.precondition, rtol=1e-3, max_iter=4000)
    assert abs(kr - k1) <= max(2, 0.01 * kr)
    assert rel(x1, xr) < 1e-6


def test_presummed_pcg_in_a_batch_equals_the_single_solve():
    from difffea_amd.newton_krylov import HipNewton
    from difffea_amd.serve import Request, Server, request_problem
    big = Request(0, 0.0, "quad4", 780, 0.4, 0.8, 0.01, n_load_steps=1)
    small = [Request(i, 0.0, "quad4" if i % 2 else "quad9", 12 + i, 0.3, 0.6, 0.05) for i in range(1, 4)]
    p, th = request_problem(big)
    solver = HipNewton(p, th.to(DEV))
    from difffea_amd.serve import reference_residual_norm
    tol = 1e-5 * reference_residual_norm(solver.sys, p, solver.device)
    assert solver.mesh.plan.n_chunks > 2048
    u_ref, info = solver.solve(1, tol, cg_max_iter=20000)
    assert info["converged"]
    big.tol = tol
    recs, summ = Server("fp64", max_slots=8).run([big] + small, realtime=False, keep_solution=True)
    assert summ["converged"] == 4
    r = {x["rid"]: x for x in recs}[0]
    assert torch.equal(r["u"], u_ref.cpu())
