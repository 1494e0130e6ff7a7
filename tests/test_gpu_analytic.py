"""Complete HIP solves against closed-form answers, as the reference's own end-to-end tests check its solver
(`tests/test_analytic.py:46-76`, `tests/test_solver_end_to_end.py:12-38`): no oracle in the loop.

* Cantilever L = 10, H = 1 under a small tip shear P (linear regime), plane strain: the tip deflection of the
  quadratic elements is within 1 % of Timoshenko beam theory, w = P L^3 / (3 E' I) + (6/5) P L / (G A) with
  E' = E / (1 - nu^2); quad4 (shear locking) approaches it from below under refinement.  The bar is the
  reference's; the plane-stress formula would be 9 % off, so it catches an E / E' mix-up.
* Finite-strain patch test: the affine motion x = F0 X imposed on the whole boundary of a distorted
  (Delaunay, plate with a hole) mesh is reproduced exactly inside, for every element type, FP64 and mixed.
"""
import pytest
import torch

from difffea_amd.boundary_conditions import boundary_edges, element_edges
from difffea_amd.problem import build_problem

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")
E, NU = 1000.0, 0.3


def _solve(p, steps, tol, precision="fp64"):
    from difffea_amd.newton_krylov import newton_krylov_solve
    th = torch.ones(p["conn"].shape[0], dtype=torch.float64, device=DEV)
    u, info = newton_krylov_solve(p, th, steps, tol, verbose=False, precision=precision, cg_max_iter=20000)
    assert info["converged"]
    return u.cpu()


def _tip_ratio(element_type, h, precision="fp64", L=10.0, H=1.0, P=-0.002):
    """Mean tip deflection of the structured cantilever over the Timoshenko value."""
    p = build_problem(dict(element_type=element_type, geometry="rectangle", geom_params=dict(width=L, height=H),
                           h=h, E=E, nu=NU, load_mode="force", load_value=(0.0, P), fixed_side="left",
                           load_side="right"))
    u = _solve(p, 2, 1e-9 * abs(P), precision)
    tip = p["nodes"][:, 0] > L - 1e-9
    w = float(u.reshape(-1, 2)[tip, 1].mean())
    E_plane = E / (1.0 - NU ** 2)
    G = E / (2.0 * (1.0 + NU))
    return w / (P * L ** 3 / (3.0 * E_plane * H ** 3 / 12.0) + 1.2 * P * L / (G * H))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("element_type", ["quad8", "quad9", "tri6"])
def test_cantilever_matches_timoshenko_beam(element_type, precision):
    assert _tip_ratio(element_type, 0.25, precision) == pytest.approx(1.0, abs=0.01)


def test_cantilever_quad4_converges_from_below_by_refinement():
    coarse, fine = _tip_ratio("quad4", 0.25), _tip_ratio("quad4", 0.125)
    assert coarse < fine < 1.0
    assert fine > 0.97


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("element_type", ["tri3", "quad4", "tri6", "quad8", "quad9"])
def test_finite_strain_patch_test(element_type, precision):
    p = build_problem(dict(element_type=element_type, mesher="unstructured", geometry="plate_hole",
                           geom_params=dict(width=2.0, height=1.0, radius=0.25), h=0.2, mesh_seed=3))
    F0 = torch.tensor([[1.10, 0.06], [-0.04, 0.95]], dtype=torch.float64)
    u_exact = (p["nodes"] @ (F0 - torch.eye(2, dtype=torch.float64)).T).reshape(-1)
    bnodes = torch.unique(boundary_edges(p["conn"], element_edges(element_type)).reshape(-1))
    fixed = torch.zeros(p["nodes"].shape[0], dtype=torch.bool)
    fixed[bnodes] = True
    fixed = fixed.repeat_interleave(2)
    assert (~fixed).any()                      # an interior to reproduce
    q = dict(p, fixed=fixed, m=torch.where(fixed, 0.0, 1.0).to(torch.float64), u_D=u_exact,
             f=torch.zeros_like(p["f"]))
    u = _solve(q, 2, 1e-10, precision)
    assert float((u - u_exact).abs().max()) < 1e-8
