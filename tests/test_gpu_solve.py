"""Complete solves and sensitivities on the GPU vs the oracle and the reference's golden vectors.
Bars: solution within 1e-10 relative (FP64), Newton counts identical; mixed precision converges to the
same FP64 residual tolerance; adjoint gradient within 1e-9 of the oracle / reference adjoint and 1e-6 of
central finite differences."""
import numpy as np
import pytest
import torch

from helpers import GOLDEN_SOLVES, cantilever, check_solution, check_trajectory, golden_max_newton, golden_problem, \
    history_array, load_golden, rel, unit_square
from oracle import fem

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")


def _solve(p, th, steps, tol, **kw):
    from difffea_amd.newton_krylov import newton_krylov_solve
    return newton_krylov_solve(p, th.to(DEV), steps, tol, verbose=False, **kw)


@pytest.mark.parametrize("et", ["quad4", "quad9"])
def test_solve_matches_oracle(et):
    p = fem.precompute_geometry(cantilever(et, h=0.25, load=(0.0, -20.0)))
    th = 0.5 + torch.rand(p["conn"].shape[0], dtype=torch.float64, generator=torch.Generator().manual_seed(1))
    tol = 1e-10 * float(torch.norm(p["f"]))
    u_ref, info_ref = fem.newton_krylov_solve(p, th, 3, tol)
    u, info = _solve(p, th, 3, tol)
    assert info["converged"] and info_ref["converged"]
    # equal Newton counts are a fair demand only when no Newton residual of either run lies near the
    # tolerance: round-off in the inexact inner solves moves a residual by tens of per cent on this strongly
    # nonlinear case (line-search steps of 1/4), so one Newton step of slack is allowed then
    near = [h["residual"] for h in info_ref["history"] + info["history"] if tol / 3 < h["residual"] < 3 * tol]
    assert abs(info["total_newton"] - info_ref["total_newton"]) <= (1 if near else 0)
    assert abs(info["total_cg"] - info_ref["total_cg"]) <= max(3, 0.03 * info_ref["total_cg"])
    assert rel(u, u_ref) < 1e-10
    assert float(torch.norm(fem.global_residual(u.cpu(), th, p))) < tol


@pytest.mark.parametrize("name", GOLDEN_SOLVES)
def test_solve_matches_reference_golden(name):
    """The reference's own trajectories, cut-backs included (after NaN steps from inverted elements, after
    max_newton, the give-up after max_cutbacks halvings, `newton_krylov.py:126-264`): convergence flag, load
    factor reached, cut-back count, Newton count, every (step, iteration, alpha) and load factor, CG counts
    within a few iterations, the callback's records, u within 1e-10."""
    z, spec = load_golden(name)
    p = golden_problem(spec)
    th = torch.ones(p["conn"].shape[0], dtype=torch.float64)
    records = []
    u, info = _solve(p, th, int(z["n_load_steps"]), float(z["tol"]), max_newton=golden_max_newton(z),
                     callback=records.append)
    assert info["converged"] == bool(z["converged"])
    assert info["n_cutbacks"] == int(z["n_cutbacks"])
    assert info["total_newton"] == int(z["total_newton"])
    if "load_factor" in z.files:
        assert info["load_factor"] == float(z["load_factor"])
        assert np.array_equal([lf for lf, _ in info["snapshots"]], z["snapshot_lf"])
    assert records == info["history"]
    check_solution(u, z)
    check_trajectory(history_array(info), z)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_nan_residual_pcg_result_equals_the_kernels(precision):
    """Where a Newton residual is NaN (inverted elements after the lift) the PCG's result is written without
    running it (``nan_pcg_result``: every entry NaN, max_iter iterations).  Running the fused PCG kernels on
    that right-hand side instead must give the same trajectory, the same cut-backs and the same bits."""
    from difffea_amd.device import HipMesh
    from difffea_amd.newton_krylov import HipNewton

    class RunPcg(HipNewton):
        def residual(self, u, lf, out=None):
            self.at_u = u                    # the last residual evaluated is the Newton state's
            return super().residual(u, lf, out)

        def linear_solve(self, rtol, max_iter, first=None, trial_at=None):
            if self._rhs_nan:
                self._rhs_nan = False
                self.sys.moduli(self.at_u)
                x, k = super().linear_solve(rtol, max_iter, first, trial_at)
                assert torch.isnan(x).all() and k == max_iter
                return x, k
            return super().linear_solve(rtol, max_iter, first, trial_at)

    z, spec = load_golden("g2_q4_nan_cutback_4")
    p = golden_problem(spec)
    th = torch.ones(p["conn"].shape[0], dtype=torch.float64, device=DEV)
    out = []
    for cls in (HipNewton, RunPcg):
        s = cls(p, th, precision=precision, mesh=HipMesh(p, DEV))
        out.append(s.solve(int(z["n_load_steps"]), float(z["tol"]), max_newton=golden_max_newton(z)))
    (u0, i0), (u1, i1) = out
    assert i0["pcg_not_run"] >= 1 and i0["n_cutbacks"] >= 1 and i0["converged"]
    assert np.array_equal(history_array(i0), history_array(i1), equal_nan=True)
    assert torch.equal(u0, u1)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("et,n", [("quad4", 30), ("quad9", 12)])
def test_trial_behind_pcg_rounds_changes_no_bit(et, n, precision):
    """The line search's first trial and its residual are enqueued behind every round of PCG kernels while the
    host reads the status; those of rounds that were not the last run while the PCG goes on and must not
    disturb it (its shared-node partials, its right-hand side): short rounds (many of them) give the same
    trajectory and bits as the solve without the early trial."""
    from difffea_amd.device import HipMesh
    from difffea_amd.newton_krylov import HipNewton
    from difffea_amd.problem import build_problem

    class NoEarlyTrial(HipNewton):
        def linear_solve(self, rtol, max_iter, first=None, trial_at=None):
            return super().linear_solve(rtol, max_iter, first, None)

    p = build_problem(dict(element_type=et, nx=n, ny=n, E=1.0, nu=0.3, load_mode="displacement",
                           load_value=(0.1, None)))
    th = torch.ones(p["conn"].shape[0], dtype=torch.float64, device=DEV)
    out = []
    for cls in (HipNewton, NoEarlyTrial):
        s = cls(p, th, precision=precision, mesh=HipMesh(p, DEV))
        s.pcg_chunk = 2
        out.append(s.solve(3, 1e-9, cg_max_iter=2000))
    (u0, i0), (u1, i1) = out
    assert i0["converged"] and i0["total_cg"] > 20
    assert np.array_equal(history_array(i0), history_array(i1))
    assert torch.equal(u0, u1)


@pytest.mark.parametrize("et", ["quad4", "quad9"])
def test_mixed_precision_solve_reaches_fp64_tolerance(et):
    p = fem.precompute_geometry(cantilever(et, h=0.25, load=(0.0, -20.0)))
    th = torch.ones(p["conn"].shape[0], dtype=torch.float64)
    tol = 1e-10 * float(torch.norm(p["f"]))
    u64, i64 = _solve(p, th, 3, tol)
    u32, i32 = _solve(p, th, 3, tol, precision="mixed")
    assert i64["converged"] and i32["converged"]
    assert float(torch.norm(fem.global_residual(u32.cpu(), th, p))) < tol
    assert rel(u32, u64) < 1e-8


def test_displacement_controlled_unit_square():
    p = unit_square("quad4", 20, value=(0.1, None))
    th = torch.ones(p["conn"].shape[0], dtype=torch.float64)
    u, info = _solve(p, th, 3, 1e-11)
    assert info["converged"]
    right = p["nodes"][:, 0] > 1 - 1e-9
    assert torch.allclose(u.cpu().reshape(-1, 2)[right, 0], torch.full((int(right.sum()),), 0.1, dtype=torch.float64))


def test_adjoint_matches_reference_golden_and_oracle():
    from difffea_amd.adjoint import compute_sensitivity
    z, spec = load_golden("g3_q4_compliance_16")
    p = golden_problem(spec)
    th = torch.from_numpy(z["theta"])
    f = p["f"].to(DEV)
    u, g = compute_sensitivity(p, th.to(DEV), lambda w: torch.dot(f, w), n_load_steps=2, tol=1e-12)
    assert rel(u, torch.from_numpy(z["u"])) < 1e-10
    assert rel(g, torch.from_numpy(z["grad"])) < 1e-9


def test_mixed_precision_adjoint_refines_to_the_fp64_gradient():
    """Mixed precision: FP32 moduli, product and inner CG, with the adjoint refined against an FP64 tangent
    (and Newton against the FP64 residual) -- the gradient reaches the reference adjoint's accuracy."""
    from difffea_amd.adjoint import compute_sensitivity
    z, spec = load_golden("g3_q4_compliance_16")
    p = golden_problem(spec)
    th = torch.from_numpy(z["theta"])
    f = p["f"].to(DEV)
    u, g = compute_sensitivity(p, th.to(DEV), lambda w: torch.dot(f, w), n_load_steps=2, tol=1e-12,
                               precision="mixed")
    assert rel(u, torch.from_numpy(z["u"])) < 1e-10
    assert rel(g, torch.from_numpy(z["grad"])) < 1e-9


def test_autograd_function_gradient_matches_finite_differences():
    from difffea_amd.adjoint import equilibrium
    p = cantilever("quad4", h=1.0, load=(50.0, 0.0))
    n_el = p["conn"].shape[0]
    th0 = 1.0 + 0.2 * torch.sin(torch.arange(n_el, dtype=torch.float64))
    f = p["f"].to(DEV)
    th = th0.to(DEV).requires_grad_(True)
    u = equilibrium(th, p, n_load_steps=2, tol=1e-12)
    L = torch.dot(f, u)
    L.backward()
    g = th.grad.detach().cpu()
    eps, gfd = 1e-6, torch.zeros(n_el, dtype=torch.float64)
    for i in range(n_el):
        tp, tm = th0.clone(), th0.clone()
        tp[i] += eps
        tm[i] -= eps
        up, _ = _solve(p, tp, 2, 1e-12)
        um, _ = _solve(p, tm, 2, 1e-12)
        gfd[i] = (float(torch.dot(f, up)) - float(torch.dot(f, um))) / (2 * eps)
    assert float(torch.norm(g - gfd) / torch.norm(gfd)) < 1e-6


def test_sequential_solves_on_fresh_meshes_release_device_memory():
    """Fifty solves, each on a freshly built problem that the caller then drops: the device memory in use
    returns to where it started (the mesh cache holds no mesh whose tensors are gone)."""
    import gc
    from difffea_amd.newton_krylov import newton_krylov_solve
    from difffea_amd.operators import cached_meshes
    p = unit_square("quad4", 6, value=(0.05, None))       # whatever a first solve sets up once, before counting
    newton_krylov_solve(p, torch.ones(p["conn"].shape[0], dtype=torch.float64, device=DEV), 1, 1e-9, verbose=False)
    del p
    gc.collect()
    torch.cuda.synchronize()
    base, base_meshes = torch.cuda.memory_allocated(), cached_meshes()
    for i in range(50):
        p = unit_square("quad4", 8 + i % 5, value=(0.05, None))
        th = torch.ones(p["conn"].shape[0], dtype=torch.float64, device=DEV)
        u, info = newton_krylov_solve(p, th, 1, 1e-9, verbose=False)
        assert info["converged"]
        del p, th, u, info
    gc.collect()
    torch.cuda.synchronize()
    assert cached_meshes() == base_meshes
    assert torch.cuda.memory_allocated() == base
