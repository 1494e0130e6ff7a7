"""Unstructured meshes on the GPU (the reference meshes arbitrary geometries with Gmsh,
`src/difffea/gmsh_mesh.py`, and feeds them to the same operators): Delaunay plates with a hole, irregular
node valence and no cell structure, for every element type.  The HIP kernels against the oracle on the same
inputs: A(v) within 1e-11 of the oracle's qp and jvp products (FP64), 1e-5 (FP32 moduli), residual, Jacobi
and block-Jacobi, fields, reactions and the adjoint gradient within 1e-11; complete solves within 1e-10,
including one read from a mesh file with clockwise elements and unused nodes; the product at ~210k
elements."""
import numpy as np
import pytest
import torch

from helpers import rel
from oracle import fem

pytestmark = pytest.mark.gpu
DEV = "cuda"
H = {"tri3": 0.03, "quad4": 0.03, "tri6": 0.06, "quad8": 0.06, "quad9": 0.06}


def _smooth_state(p, h, seed):
    """Smooth finite deformation (stretch, shear waves: |grad u| <= 0.1 on any element, however small) plus
    N(0, (1e-3 h)^2) per dof, and a random probe direction."""
    g = torch.Generator().manual_seed(seed)
    X = p["nodes"]
    u = torch.stack([0.05 * X[:, 0] + 0.03 * torch.sin(3 * X[:, 1]),
                     -0.02 * X[:, 1] + 0.02 * torch.cos(2 * X[:, 0])], 1).reshape(-1)
    u = u + 1e-3 * h * torch.randn(u.numel(), dtype=torch.float64, generator=g)
    return u, torch.randn(u.numel(), dtype=torch.float64, generator=g)


def _plate(et, h, **kw):
    from difffea_amd.problem import build_problem
    spec = dict(element_type=et, mesher="unstructured", geometry="plate_hole",
                geom_params=dict(width=1.0, height=1.0, radius=0.2), h=h, E=1.0, nu=0.3,
                load_mode="force", load_value=(0.1, 0.02), fixed_side="left", load_side="right")
    spec.update(kw)
    return build_problem(spec)


@pytest.mark.parametrize("et", list(H))
def test_unstructured_kernels_match_oracle(et):
    from difffea_amd.operators import compute_jacobi_diagonal, global_residual, make_system, make_tangent_operator
    from difffea_amd.postprocess import compute_fields, support_reactions
    p = fem.precompute_geometry(_plate(et, H[et]))
    print(f"case {et} h={H[et]} n_el={p['conn'].shape[0]}", flush=True)
    u, v = _smooth_state(p, H[et], 1)
    th = 0.5 + torch.rand(p["conn"].shape[0], dtype=torch.float64, generator=torch.Generator().manual_seed(2))
    assert torch.isfinite(fem.global_residual(u, th, p)).all()
    ud, thd, vd = u.to(DEV), th.to(DEV), v.to(DEV)
    assert rel(global_residual(ud, thd, p, 0.6), fem.global_residual(u, th, p, 0.6)) < 1e-11
    y = make_tangent_operator("hip", ud, thd, p)(vd)
    assert rel(y, fem.make_A_operator_qp(u, th, p)(v)) < 1e-11
    assert rel(y, fem.make_A_operator(u, th, p)(v)) < 1e-11
    y32 = make_tangent_operator("hip", ud, thd, p, precision="mixed")(vd.float())
    assert rel(y32, y) < 1e-5
    assert rel(compute_jacobi_diagonal(ud, thd, p), fem.compute_jacobi_diagonal(u, th, p)) < 1e-11
    s = make_system(ud, thd, p, preconditioner="block_jacobi")
    s.moduli(ud)
    Binv = torch.linalg.inv(fem.block_jacobi_blocks(u, th, p))
    nd = s.mesh.n_dof
    got = torch.stack([s.diag[:nd].reshape(-1, 2)[:, 0], s.diag[nd:], s.diag[:nd].reshape(-1, 2)[:, 1]], 1).cpu()
    assert rel(got, torch.stack([Binv[:, 0, 0], Binv[:, 0, 1], Binv[:, 1, 1]], 1)) < 1e-10
    out, ref = compute_fields(ud, thd, p), fem.compute_fields(u, th, p)
    for k in ("J", "sxx", "syy", "sxy", "mises", "energy_density", "J_n", "mises_n", "energy_density_n"):
        assert rel(out[k], ref[k]) < 1e-11, k
    assert rel(support_reactions(ud, thd, p), fem.support_reactions(u, th, p)) < 1e-11
    lam = torch.randn(u.shape[0], dtype=torch.float64, generator=torch.Generator().manual_seed(3))
    assert rel(make_system(ud, thd, p).adjoint_grad(ud, lam.to(DEV)), fem.compute_implicit_vjp(u, th, p, lam)) < 1e-11


@pytest.mark.parametrize("et", ["quad4", "tri6"])
def test_unstructured_solve_from_a_mesh_file(et, tmp_path):
    """A plate with a hole written as a mesher would (shuffled node ids, unused points, half the elements
    clockwise), loaded through mesher='npz', solved on the GPU and by the oracle."""
    from difffea_amd.mesh import create_unstructured_plate_hole
    from difffea_amd.newton_krylov import newton_krylov_solve
    from difffea_amd.postprocess import support_reactions
    nodes, conn = create_unstructured_plate_hole(et, h=0.08, seed=5)
    rng = np.random.default_rng(6)
    N = nodes.shape[0]
    perm = rng.permutation(N + 3)
    xy = np.zeros((N + 3, 2))
    xy[perm[:N]] = nodes.numpy()
    xy[perm[N:]] = 7.0
    c = conn.numpy().copy()
    flip = rng.random(c.shape[0]) < 0.5
    c[flip] = c[flip][:, {"quad4": [0, 3, 2, 1], "tri6": [0, 2, 1, 5, 4, 3]}[et]]
    np.savez(tmp_path / "mesh.npz", xy=xy, conn=perm[c])
    p = fem.precompute_geometry(_plate(et, 0.08, mesher="npz", mesh_file=str(tmp_path / "mesh.npz")))
    assert float(p["w_detJ"].min()) > 0
    th = 0.5 + torch.rand(p["conn"].shape[0], dtype=torch.float64, generator=torch.Generator().manual_seed(7))
    tol = 1e-10 * float(torch.norm(p["f"]))
    u_ref, info_ref = fem.newton_krylov_solve(p, th, 3, tol, cg_max_iter=5000)
    u, info = newton_krylov_solve(p, th.to(DEV), 3, tol, verbose=False, cg_max_iter=5000)
    assert info["converged"] and info_ref["converged"]
    assert abs(info["total_newton"] - info_ref["total_newton"]) <= 1
    assert rel(u, u_ref) < 1e-10
    r = support_reactions(u, th.to(DEV), p).cpu()
    assert torch.allclose(r + p["f"].reshape(-1, 2).sum(0), torch.zeros(2, dtype=torch.float64), atol=1e-9)
    um, im = newton_krylov_solve(p, th.to(DEV), 3, tol, verbose=False, precision="mixed", cg_max_iter=5000)
    assert im["converged"] and rel(um, u_ref) < 1e-8


def test_unstructured_product_at_scale():
    """About 210k quadrilaterals with valences 1..8: FP64 A(v) within 1e-11 of the oracle's qp product."""
    from difffea_amd.operators import make_tangent_operator
    p = _plate("quad4", 0.0025)
    n_el = p["conn"].shape[0]
    print(f"case quad4 h=0.0025 n_el={n_el}", flush=True)
    assert n_el > 200_000
    p = fem.precompute_geometry(p)
    u, v = _smooth_state(p, 0.0025, 4)
    th = torch.ones(n_el, dtype=torch.float64)
    C, dN = fem.qp_setup(u, th, p)
    ref = fem.make_A_operator_qp(u, th, p, C=C)(v)
    y = make_tangent_operator("hip", u.to(DEV), th.to(DEV), p)(v.to(DEV))
    assert rel(y, ref) < 1e-11
