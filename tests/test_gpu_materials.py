"""The material library in the HIP kernels (hand-derived closed forms of `materials.py:65-162` of the
reference) against the oracle's automatic differentiation of the same energies: residual, moduli through
A(v), Jacobi diagonal, block-Jacobi blocks, post-processing fields and the adjoint gradient within 1e-11
relative (FP64); complete solves within 1e-10; the neo-Hookean-only kernels refuse other materials with NaN;
the fibre material's compliance gradient against central differences of complete solves (1e-6)."""
import pytest
import torch

from helpers import cantilever, rel, state, unit_square
from oracle import fem

pytestmark = pytest.mark.gpu
DEV = "cuda"
CASES = [("mooney_rivlin", None), ("saint_venant", None), ("demiray", None), ("fiber_reinforced", None),
         ("fiber_reinforced", {"d": 0.2, "beta_deg": 60.0, "second_family": 0.0}), ("neo_hookean", None)]


def _setup(name, vals, et, n, seed=0):
    p = fem.precompute_geometry(unit_square(et, n, material=name, material_params=vals))
    amp = 0.06 if et == "quad4" else 0.03        # |F - I| up to about 0.45, J > 0.6
    u, v = state(p, seed, scale=amp / n)
    th = 0.5 + torch.rand(p["conn"].shape[0], dtype=torch.float64, generator=torch.Generator().manual_seed(seed + 7))
    p = dict(p, f=torch.randn(p["f"].shape[0], dtype=torch.float64, generator=torch.Generator().manual_seed(5)))
    return p, u, v, th


@pytest.mark.parametrize("et,n", [("quad4", 9), ("quad9", 5)])
@pytest.mark.parametrize("name,vals", CASES)
def test_material_kernels_match_oracle(name, vals, et, n):
    from difffea_amd.operators import compute_jacobi_diagonal, global_residual, make_system, make_tangent_operator
    from difffea_amd.postprocess import compute_fields
    p, u, v, th = _setup(name, vals, et, n, seed=1)
    ud, thd = u.to(DEV), th.to(DEV)
    assert rel(global_residual(ud, thd, p, 0.7), fem.global_residual(u, th, p, 0.7)) < 1e-11
    A = make_tangent_operator("hip", ud, thd, p)
    assert rel(A(v.to(DEV)), fem.make_A_operator_qp(u, th, p)(v)) < 1e-11
    assert rel(A(v.to(DEV)), fem.make_A_operator(u, th, p)(v)) < 1e-11
    assert rel(compute_jacobi_diagonal(ud, thd, p), fem.compute_jacobi_diagonal(u, th, p)) < 1e-11
    s = make_system(ud, thd, p, preconditioner="block_jacobi")
    s.moduli(ud)
    Binv = torch.linalg.inv(fem.block_jacobi_blocks(u, th, p))
    nd = s.mesh.n_dof
    got = torch.stack([s.diag[:nd].reshape(-1, 2)[:, 0], s.diag[nd:], s.diag[:nd].reshape(-1, 2)[:, 1]], 1).cpu()
    want = torch.stack([Binv[:, 0, 0], Binv[:, 0, 1], Binv[:, 1, 1]], 1)
    assert rel(got, want) < 1e-10
    out, ref = compute_fields(ud, thd, p), fem.compute_fields(u, th, p)
    for k in ("J", "sxx", "syy", "sxy", "mises", "energy_density", "mises_n", "energy_density_n"):
        assert rel(out[k], ref[k]) < 1e-11, k
    lam = torch.randn(u.shape[0], dtype=torch.float64, generator=torch.Generator().manual_seed(3))
    g = make_system(ud, thd, p).adjoint_grad(ud, lam.to(DEV))
    assert rel(g, fem.compute_implicit_vjp(u, th, p, lam)) < 1e-11


def test_halo_plan_and_mixed_precision_with_a_library_material():
    from difffea_amd.device import HipMesh, HipSystem
    from difffea_amd.operators import _check_material
    p, u, v, th = _setup("fiber_reinforced", {"d": 0.1}, "quad4", 37, seed=2)
    ref = fem.make_A_operator_qp(u, th, p)(v)
    for mode, prec, tol in (("halo", "fp64", 1e-11), ("partials", "mixed", 1e-5)):
        mesh = HipMesh(p, DEV, plan_mode=mode)
        s = HipSystem(mesh, th, _check_material(p), p["f"], prec)
        s.moduli(u.to(DEV))
        assert rel(s.apply(v.to(DEV)), ref) < tol, (mode, prec)
        assert rel(s.residual(u.to(DEV), lf=0.4), fem.global_residual(u, th, p, 0.4)) < 1e-11


@pytest.mark.parametrize("name", ["mooney_rivlin", "saint_venant", "demiray", "fiber_reinforced"])
def test_library_material_solve_matches_oracle(name):
    from difffea_amd.newton_krylov import newton_krylov_solve
    from difffea_amd.postprocess import support_reactions
    p = fem.precompute_geometry(cantilever("quad4", h=0.25, load=(0.0, -15.0), material=name))
    th = 0.5 + torch.rand(p["conn"].shape[0], dtype=torch.float64, generator=torch.Generator().manual_seed(1))
    tol = 1e-10 * float(torch.norm(p["f"]))
    u_ref, info_ref = fem.newton_krylov_solve(p, th, 3, tol)
    u, info = newton_krylov_solve(p, th.to(DEV), 3, tol, verbose=False)
    assert info["converged"] and info_ref["converged"]
    near = [h["residual"] for h in info_ref["history"] + info["history"] if tol / 3 < h["residual"] < 3 * tol]
    assert abs(info["total_newton"] - info_ref["total_newton"]) <= (1 if near else 0)
    assert rel(u, u_ref) < 1e-10
    r = support_reactions(u, th.to(DEV), p).cpu()
    assert torch.allclose(r + p["f"].reshape(-1, 2).sum(0), torch.zeros(2, dtype=torch.float64), atol=1e-7)
    um, im = newton_krylov_solve(p, th.to(DEV), 3, tol, verbose=False, precision="mixed")
    assert im["converged"] and rel(um, u_ref) < 1e-8


def test_neo_hookean_only_kernels_refuse_other_materials():
    """Without DFEA_MATLIB in the element kind, a Mooney-Rivlin descriptor gets NaN from the material
    kernels instead of another energy's numbers."""
    from difffea_amd import _lib
    from difffea_amd.device import _ptr, _stream
    from difffea_amd.operators import make_system
    p, u, v, th = _setup("mooney_rivlin", None, "quad4", 6)
    s = make_system(u.to(DEV), th.to(DEV), p)
    assert s.kind & _lib.MATLIB
    m = s.mesh
    out = torch.empty(m.n_dof, dtype=torch.float64, device=DEV)
    ud = u.to(DEV)
    _lib.check(s.lib.dfea_residual(_ptr(s.table), _ptr(m.items), m.plan.n_chunks, _ptr(m.fin_items), m.n_finblk,
                                   _ptr(ud), _ptr(out), 1.0, m.kind, _stream()), "dfea_residual")
    free = p["m"].to(DEV) > 0
    assert bool(torch.isnan(out[free]).all())
    good = s.residual(u.to(DEV), lf=1.0)
    assert bool(torch.isfinite(good).all())


def test_adjoint_sensitivity_with_fibre_material_matches_finite_differences():
    """The reference's fibre-material sensitivity check (`tests/test_materials.py:101-114`) through the HIP path:
    compliance gradient from the adjoint solve and the gradient kernel against central differences of complete
    HIP solves (step 1e-6) within 1e-6 relative."""
    from difffea_amd.adjoint import compute_sensitivity
    from difffea_amd.newton_krylov import solve_newton_krylov
    p = cantilever("quad4", h=1.0, load=(60.0, 0.0), material="fiber_reinforced")
    n_el = p["conn"].shape[0]
    theta = (1.0 + 0.2 * torch.sin(torch.arange(n_el, dtype=torch.float64))).to(DEV)
    f = p["f"].to(DEV)

    def loss(u):
        return torch.dot(f, u)

    _, g = compute_sensitivity(p, theta, loss, n_load_steps=2, tol=1e-12)
    eps, g_fd = 1e-6, torch.zeros(n_el, dtype=torch.float64)
    for i in range(n_el):
        tp, tm = theta.clone(), theta.clone()
        tp[i] += eps
        tm[i] -= eps
        g_fd[i] = (float(loss(solve_newton_krylov(p, tp, 2, 1e-12, verbose=False))) -
                   float(loss(solve_newton_krylov(p, tm, 2, 1e-12, verbose=False)))) / (2 * eps)
    assert float(torch.norm(g.cpu() - g_fd) / torch.norm(g_fd)) < 1e-6
