"""HIP kernels vs the CPU oracle on the same inputs (FP64 bar: 1e-11 relative, the reference's own
cross-variant tolerance, test_tangent_operators.py:31-36)."""
import pytest
import torch

from helpers import cantilever, rel, state, stretch_state, unit_square
from oracle import fem

pytestmark = pytest.mark.gpu
DEV = "cuda"
CASES = [("quad4", 8), ("quad9", 4), ("quad4", 37), ("quad9", 19)]


def _setup(et, n, seed=0, nonuniform=False, amp=None):
    p = fem.precompute_geometry(unit_square(et, n))
    # 0.02 at n = 8 for Q4 (the reference tests' amplitude); Q9 nodes are twice as dense, so a quarter of it
    amp = amp if amp is not None else (0.16 if et == "quad4" else 0.04)
    u, v = state(p, seed, scale=amp / n)
    assert torch.isfinite(fem.global_residual(u, torch.ones(p["conn"].shape[0], dtype=torch.float64), p)).all()
    th = torch.ones(p["conn"].shape[0], dtype=torch.float64)
    if nonuniform:
        th = 0.5 + torch.rand(th.shape[0], dtype=torch.float64, generator=torch.Generator().manual_seed(seed + 7))
    return p, u, v, th


@pytest.mark.parametrize("et,n", CASES)
def test_apply_fp64_matches_oracle(et, n):
    from difffea_amd.operators import make_tangent_operator
    p, u, v, th = _setup(et, n, nonuniform=True)
    ref_qp = fem.make_A_operator_qp(u, th, p)(v)
    ref_jvp = fem.make_A_operator(u, th, p)(v)
    A = make_tangent_operator("hip", u.to(DEV), th.to(DEV), p)
    y = A(v.to(DEV))
    assert rel(y, ref_qp) < 1e-11
    assert rel(y, ref_jvp) < 1e-11
    assert rel(y, fem.make_A_operator_linearize(u, th, p)(v)) < 1e-11


@pytest.mark.parametrize("et,n", CASES)
def test_diagonal_and_residual_match_oracle(et, n):
    from difffea_amd.operators import compute_jacobi_diagonal, global_residual
    p, u, v, th = _setup(et, n, seed=3, nonuniform=True)
    p = dict(p, f=torch.randn(p["f"].shape[0], dtype=torch.float64, generator=torch.Generator().manual_seed(5)))
    d = compute_jacobi_diagonal(u.to(DEV), th.to(DEV), p)
    assert rel(d, fem.compute_jacobi_diagonal(u, th, p)) < 1e-11
    R = global_residual(u.to(DEV), th.to(DEV), p, 0.7)
    assert rel(R, fem.global_residual(u, th, p, 0.7)) < 1e-11


@pytest.mark.parametrize("et,n", [("quad4", 8), ("quad9", 4)])
def test_apply_is_bitwise_repeatable_and_symmetric(et, n):
    from difffea_amd.operators import make_tangent_operator
    p, u, v, th = _setup(et, n, seed=4)
    A = make_tangent_operator("hip", u.to(DEV), th.to(DEV), p)
    y1, y2 = A(v.to(DEV)), A(v.to(DEV))
    assert torch.equal(y1, y2)
    w = torch.randn_like(v)
    a, b = float(w.to(DEV) @ A(v.to(DEV))), float(v.to(DEV) @ A(w.to(DEV)))
    assert a == pytest.approx(b, rel=1e-10)


@pytest.mark.parametrize("et,n", [("quad4", 8), ("quad9", 4)])
def test_apply_mixed_precision(et, n):
    from difffea_amd.operators import make_tangent_operator
    p, u, v, th = _setup(et, n, seed=6)
    ref = fem.make_A_operator_qp(u, th, p)(v)
    y = make_tangent_operator("hip", u.to(DEV), th.to(DEV), p, precision="mixed")(v.to(DEV).float())
    assert y.dtype == torch.float32
    assert rel(y, ref) < 1e-5


@pytest.mark.parametrize("et,n", [("quad4", 12), ("quad9", 6)])
def test_fused_pcg_matches_oracle_pcg(et, n):
    from difffea_amd.operators import make_system
    p, u, v, th = _setup(et, n, seed=8, amp=0.01)      # near the reference state: SPD tangent
    b = v * p["m"]
    A = fem.make_A_operator_qp(u, th, p)
    M = fem.make_jacobi_preconditioner(u, th, p)
    x_ref, k_ref = fem.pcg(A, b, M, rtol=1e-10, max_iter=2000)
    s = make_system(u.to(DEV), th.to(DEV), p)
    s.moduli(u.to(DEV))
    x, k, st = s.pcg_solve(b.to(DEV), rtol=1e-10, max_iter=2000)
    assert st == 1
    assert abs(k - k_ref) <= 1
    assert rel(x, x_ref) < 1e-9


def test_pcg_zero_rhs_and_iteration_cap():
    from difffea_amd.operators import make_system
    p, u, v, th = _setup("quad4", 8, seed=9, amp=0.01)
    s = make_system(u.to(DEV), th.to(DEV), p)
    s.moduli(u.to(DEV))
    x, k, st = s.pcg_solve(torch.zeros_like(v).to(DEV), rtol=1e-8)
    assert (st, k) == (1, 0) and float(x.abs().max()) == 0.0
    b = v * p["m"]
    x, k, st = s.pcg_solve(b.to(DEV), rtol=1e-14, max_iter=5)
    x_ref, k_ref = fem.pcg(fem.make_A_operator_qp(u, th, p), b, fem.make_jacobi_preconditioner(u, th, p),
                           rtol=1e-14, max_iter=5)
    assert (st, k, k_ref) == (3, 5, 5)
    assert rel(x, x_ref) < 1e-11


@pytest.mark.parametrize("et,n", [("quad4", 8), ("quad9", 4)])
def test_adjoint_gradient_kernel_matches_oracle_vjp(et, n):
    from difffea_amd.operators import make_system
    p, u, v, th = _setup(et, n, seed=10, nonuniform=True)
    lam = v
    ref = fem.compute_implicit_vjp(u, th, p, lam)
    s = make_system(u.to(DEV), th.to(DEV), p)
    g = s.adjoint_grad(u.to(DEV), lam.to(DEV))
    assert rel(g, ref) < 1e-11


@pytest.mark.parametrize("et,n", [("quad4", 400), ("quad9", 200), ("quad4", 566)])
def test_apply_at_full_size(et, n):
    """The bench workloads: 400x400 Q4 and 200x200 Q9 with the stretched, noisy state; 566x566 Q4 is a launch of
    1,252 chunks, beyond the workgroup slots of the issue-ahead variant (the element stream starts before the
    gather there)."""
    from difffea_amd.operators import make_tangent_operator
    p = fem.precompute_geometry(unit_square(et, n))
    u = stretch_state(p, 1.0 / n, seed=0)
    v = torch.randn(u.shape[0], dtype=torch.float64, generator=torch.Generator().manual_seed(1))
    th = torch.ones(p["conn"].shape[0], dtype=torch.float64)
    ref = fem.make_A_operator_qp(u, th, p)(v)
    y = make_tangent_operator("hip", u.to(DEV), th.to(DEV), p)(v.to(DEV))
    assert rel(y, ref) < 1e-11
    y32 = make_tangent_operator("hip", u.to(DEV), th.to(DEV), p, precision="mixed")(v.to(DEV).float())
    assert rel(y32, ref) < 1e-5


@pytest.mark.parametrize("et,n", [("quad4", 37), ("quad9", 19)])
def test_pcg_graph_replay_equals_direct_launches(et, n):
    """Chunks replayed from captured HIP graphs give the same bits as direct launches, replay after replay."""
    from difffea_amd.operators import make_system
    p, u, v, th = _setup(et, n, seed=11, amp=0.01)
    b = (v * p["m"]).to(DEV)
    s = make_system(u.to(DEV), th.to(DEV), p)
    s.moduli(u.to(DEV))
    x1, k1, st1 = s.pcg_solve(b, rtol=1e-9, max_iter=3000, graph=False)
    x1 = x1.clone()
    for _ in range(3):
        x2, k2, st2 = s.pcg_solve(b, rtol=1e-9, max_iter=3000, graph=True)
        torch.cuda.synchronize()
        assert (k1, st1) == (k2, st2)
        assert torch.equal(x1, x2)


@pytest.mark.parametrize("n", [37, 64])
def test_halo_plan_matches_oracle_and_partials_plan(n):
    """Quad4 halo plans (no shared nodes, halo elements evaluated by the owner chunk): A(v), the Jacobi
    diagonal, the residual and the fused PCG agree with the oracle and with the default partials plan."""
    from difffea_amd.device import HipMesh, HipSystem
    from difffea_amd.operators import _check_material
    p, u, v, th = _setup("quad4", n, seed=12, nonuniform=True, amp=0.01)
    mh = HipMesh(p, DEV, plan_mode="halo")
    assert mh.plan.mode == "halo" and mh.plan.n_shared == 0 and len(mh.plan.hel) > 0
    sh = HipSystem(mh, th, _check_material(p), p["f"], "fp64")
    sh.moduli(u.to(DEV))
    assert rel(sh.apply(v.to(DEV)), fem.make_A_operator_qp(u, th, p)(v)) < 1e-11
    assert rel(sh.diag, fem.compute_jacobi_diagonal(u, th, p)) < 1e-11
    assert rel(sh.residual(u.to(DEV), lf=0.3), fem.global_residual(u, th, p, 0.3)) < 1e-11
    b = (v * p["m"]).to(DEV)
    x_h, k_h, st_h = sh.pcg_solve(b, rtol=1e-9, max_iter=3000)
    mp = HipMesh(p, DEV, plan_mode="partials")
    sp = HipSystem(mp, th, _check_material(p), p["f"], "fp64")
    sp.moduli(u.to(DEV))
    x_p, k_p, st_p = sp.pcg_solve(b, rtol=1e-9, max_iter=3000)
    assert st_h == st_p == 1 and abs(k_h - k_p) <= 1
    assert rel(x_h, x_p) < 1e-8


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_pcg_negative_curvature_at_first_iteration_returns_z(prec):
    """theta = -1 makes the tangent negative definite: p.Ap < 0 at k = 0, so PCG stops with status
    'curvature' after 0 iterations and returns z = M^-1 b, as the reference's pcg does."""
    from difffea_amd.operators import make_system
    p, u, v, th = _setup("quad4", 8, seed=14, amp=0.01)
    th = -th
    b = v * p["m"]
    x_ref, k_ref = fem.pcg(fem.make_A_operator_qp(u, th, p), b, fem.make_jacobi_preconditioner(u, th, p),
                           rtol=1e-8, max_iter=100)
    s = make_system(u.to(DEV), th.to(DEV), p, prec)
    s.moduli(u.to(DEV))
    x, k, st = s.pcg_solve(b.to(DEV).to(s.vdt), rtol=1e-8, max_iter=100)
    assert (st, k, k_ref) == (2, 0, 0)
    assert rel(x, x_ref) < (1e-12 if prec == "fp64" else 1e-6)


def test_apply_properties_at_the_largest_served_size():
    """Side-512 quad9 (262,144 elements, 2,101,250 dofs; the largest mesh of the served stream), no oracle at
    this size: the product is symmetric (a.A b = b.A a), linear, repeatable bit for bit, the identity on
    fixed dofs, and its FP32-moduli form agrees with FP64 to single precision."""
    from difffea_amd.operators import make_tangent_operator
    p = unit_square("quad9", 512)
    n = p["nodes"].shape[0] * 2
    u = stretch_state(p, 1.0 / 512, seed=3)
    g = torch.Generator().manual_seed(4)
    a = torch.randn(n, dtype=torch.float64, generator=g).to(DEV)
    b = torch.randn(n, dtype=torch.float64, generator=g).to(DEV)
    th = torch.ones(p["conn"].shape[0], dtype=torch.float64)
    A = make_tangent_operator("hip", u.to(DEV), th.to(DEV), p)
    Aa, Ab = A(a).clone(), A(b).clone()
    assert torch.equal(A(a), Aa)
    s1, s2 = float(torch.dot(b, Aa)), float(torch.dot(a, Ab))
    assert abs(s1 - s2) <= 1e-12 * float(torch.dot(Aa.abs(), b.abs()))
    lin = A(2.0 * a - 0.5 * b)
    assert rel(lin, 2.0 * Aa - 0.5 * Ab) < 1e-13
    fixed = p["fixed"].to(DEV)
    assert torch.equal(Aa[fixed], a[fixed])
    A32 = make_tangent_operator("hip", u.to(DEV), th.to(DEV), p, precision="mixed")
    assert rel(A32(a.float()), Aa) < 1e-5


@pytest.mark.parametrize("et,n,mode", [("quad4", 400, "auto"), ("quad9", 64, "auto"), ("tri3", 90, "auto"),
                                       ("quad4", 37, "halo")])
def test_apply_host_equals_apply_bitwise(et, n, mode):
    """dfea_apply_host (the single-solve product whose shared-node finish takes its pointers as kernel arguments)
    gives the same bits as dfea_apply, in both precisions, on partials (tail-balanced included) and halo plans."""
    import ctypes
    from difffea_amd import _lib
    from difffea_amd.device import HipMesh, HipSystem
    from difffea_amd.materials import material_spec
    p = unit_square(et, n)
    u = stretch_state(p, 1.0 / n, seed=9)
    th = torch.ones(p["conn"].shape[0], dtype=torch.float64)
    m = HipMesh(p, DEV, plan_mode=mode)
    assert m.plan.mode == ("halo" if mode == "halo" else "partials")
    for precision in ("fp64", "mixed"):
        s = HipSystem(m, th, material_spec(p), p["f"], precision)
        s.moduli(u.to(DEV))
        v = torch.randn(u.shape[0], dtype=torch.float64, generator=torch.Generator().manual_seed(2)).to(DEV).to(s.vdt)
        y1, y2 = torch.empty_like(v), torch.full_like(v, float("nan"))
        _lib.check(s.lib.dfea_apply(ctypes.c_void_p(s.table.data_ptr()), None, m.plan.n_chunks, None, m.n_finblk,
                                    ctypes.c_void_p(v.data_ptr()), ctypes.c_void_p(y1.data_ptr()), s.kind, s.prec_id,
                                    ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "dfea_apply")
        s.apply(v, y2)
        torch.cuda.synchronize()
        assert torch.equal(y1, y2)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("et,n", [("quad9", 14), ("quad8", 16), ("tri6", 11)])
def test_wide_pcg_workgroups_agree_with_oracle_and_default_groups(et, n, precision):
    """Quadratic meshes of at most 256 chunks carry DFEA_WIDE: their PCG kernel gives each element more
    Gauss-point groups (192 / 256 threads per chunk).  Against the oracle's PCG and against the same mesh with
    the flag cleared (the default 128 / 192 threads): same iteration count within one, solutions within the
    solve tolerance."""
    from difffea_amd import _lib
    from difffea_amd.device import HipMesh, HipSystem
    from difffea_amd.materials import material_spec
    p, u, v, th = _setup(et, n, seed=13, nonuniform=True, amp=0.01)
    b = v * p["m"]
    x_ref, k_ref = fem.pcg(fem.make_A_operator_qp(u, th, p), b, fem.make_jacobi_preconditioner(u, th, p),
                           rtol=1e-10, max_iter=3000)
    out = {}
    for wide in (True, False):
        m = HipMesh(p, DEV)
        assert m.kind & _lib.WIDE
        if not wide:
            m.kind &= ~_lib.WIDE
        s = HipSystem(m, th, material_spec(p), p["f"], precision)
        s.moduli(u.to(DEV))
        x, k, st = s.pcg_solve(b.to(DEV).to(s.vdt), rtol=1e-10 if precision == "fp64" else 1e-6, max_iter=3000)
        assert st == _lib.PCG_CONVERGED
        if precision == "fp64":
            assert abs(k - k_ref) <= 1
        assert rel(x, x_ref) < (1e-8 if precision == "fp64" else 1e-4)
        out[wide] = (x.double().cpu(), k)
    assert rel(out[True][0], out[False][0]) < (1e-9 if precision == "fp64" else 1e-4)


@pytest.mark.parametrize("et,n,precision", [("quad4", 700, "fp64"), ("quad9", 512, "mixed")])
def test_pcg_with_streamed_element_data_matches_textbook_pcg(et, n, precision):
    """Launches whose element data exceed the Infinity Cache (700x700 quad4 FP64: 282 MB of moduli and gradients;
    side-512 quad9 FP32: 264 MB) run the PCG kernel with nontemporal element loads.  No oracle at this size: after
    exactly 60 iterations (the cap) the fused kernel's iterate equals the textbook PCG built from the hip product,
    the system's Jacobi diagonal and torch vector operations -- an independent arrangement of the same method."""
    from difffea_amd.newton_krylov import pcg
    from difffea_amd.operators import make_system
    p = unit_square(et, n)
    u = stretch_state(p, 1.0 / n, seed=5)
    th = torch.ones(p["conn"].shape[0], dtype=torch.float64)
    s = make_system(u.to(DEV), th.to(DEV), p, precision)
    m = s.mesh
    elem_bytes = m.n_el * m.nq * (10 + 2 * m.nen) * (8 if precision == "fp64" else 4)
    assert elem_bytes > 240 * (1 << 20)
    s.moduli(u.to(DEV))
    b = (torch.randn(u.shape[0], dtype=torch.float64, generator=torch.Generator().manual_seed(6)) * p["m"]).to(DEV)
    x, k, st = s.pcg_solve(b, rtol=1e-14, max_iter=60)
    assert (st, k) == (3, 60)
    bb = b.to(s.vdt)
    x_ref, k_ref = pcg(lambda w: s.apply(w), bb, s.precondition, rtol=1e-14, max_iter=60)
    assert k_ref == 60
    assert rel(x, x_ref) < (1e-10 if precision == "fp64" else 1e-3)


@pytest.mark.parametrize("et,n", [("quad4", 37), ("quad9", 19), ("tri6", 13)])
def test_precompute_geometry_matches_oracle(et, n):
    """dN_dX and w_detJ of the geometry kernel, read back in the original element order, against the
    oracle's Jacobian inverse (`operators.py:85-120`); get_geometry returns cached entries unchanged."""
    from difffea_amd.operators import get_geometry, precompute_geometry
    p = unit_square(et, n)
    ref = fem.precompute_geometry(p)
    g = precompute_geometry(p)
    assert g["dN_dX"].shape == ref["dN_dX"].shape and g["w_detJ"].shape == ref["w_detJ"].shape
    # closed-form 2x2 inverse and determinant against LU: a few units in the last place
    assert rel(g["dN_dX"], ref["dN_dX"]) < 1e-13 and rel(g["w_detJ"], ref["w_detJ"]) < 1e-13
    dN, w = get_geometry(g)
    assert dN is g["dN_dX"] and w is g["w_detJ"]


def _same_nan(a, b):
    a, b = a.double().cpu(), b.double().cpu()
    nan = torch.isnan(b)
    assert torch.equal(torch.isnan(a), nan)
    assert float(b[~nan].abs().max()) > 0.0
    return rel(a[~nan], b[~nan])


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_inverted_elements_give_the_reference_nans(prec):
    """A lift that pushes the right edge of a 4x4 square 0.5 to the left inverts the right-hand column:
    log J is NaN there (`constitutive_relations.py:46`).  The reference keeps the NaN through the residual's
    mask (`operators.py:220`), the Jacobi clamp (`operators.py:555`) and the product's pivot
    (`operators.py:380`), and its PCG runs a NaN curvature on to the iteration cap (`newton_krylov.py:70-71`);
    the HIP path has NaN in the same entries, the oracle's values elsewhere, and runs to the cap too."""
    from difffea_amd.operators import compute_jacobi_diagonal, global_residual, make_system, make_tangent_operator
    p = fem.precompute_geometry(unit_square("quad4", 4, value=(-0.5, None)))
    th = torch.ones(p["conn"].shape[0], dtype=torch.float64)
    g = torch.Generator().manual_seed(2)
    u = torch.where(p["fixed"], p["u_D"], 0.01 * torch.randn(p["u_D"].shape[0], dtype=torch.float64, generator=g))
    v = torch.randn(u.shape[0], dtype=torch.float64, generator=g)
    Rref = fem.global_residual(u, th, p)
    assert torch.isnan(Rref).any() and not torch.isnan(Rref).all()
    tol = 1e-11 if prec == "fp64" else 1e-5
    assert _same_nan(global_residual(u.to(DEV), th.to(DEV), p), Rref) < 1e-11
    assert _same_nan(compute_jacobi_diagonal(u.to(DEV), th.to(DEV), p), fem.compute_jacobi_diagonal(u, th, p)) < 1e-11
    A = make_tangent_operator("hip", u.to(DEV), th.to(DEV), p, precision=prec)
    assert _same_nan(A(v.to(DEV)), fem.make_A_operator_qp(u, th, p)(v)) < tol
    s = make_system(u.to(DEV), th.to(DEV), p, prec)
    s.moduli(u.to(DEV))
    _, k_ref = fem.pcg(fem.make_A_operator_qp(u, th, p), -Rref, fem.make_jacobi_preconditioner(u, th, p),
                       rtol=0.5, max_iter=30)
    x, k, st = s.pcg_solve((-Rref).to(DEV).to(s.vdt), rtol=0.5, max_iter=30)
    assert (st, k, k_ref) == (3, 30, 30)
    assert torch.isnan(x).all()           # the result `newton_krylov.nan_pcg_result` writes without a run


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_small_and_large_chunk_plans_agree_with_oracle(precision):
    """The same quad4 mesh through small-chunk plans (64-element chunks, 4 threads per element in the product
    and PCG) and the default 256-element plans: product, residual, diagonal and PCG against the oracle and
    against each other."""
    from difffea_amd import _lib
    from difffea_amd.device import HipMesh, HipSystem
    from difffea_amd.materials import material_spec
    p, u, v, th = _setup("quad4", 60, seed=11, nonuniform=True, amp=0.02)     # a positive definite tangent
    tol = 1e-11 if precision == "fp64" else 1e-5
    ref = fem.make_A_operator_qp(u, th, p)(v)
    b = v * p["m"]
    x_ref, k_ref = fem.pcg(fem.make_A_operator_qp(u, th, p), b, fem.make_jacobi_preconditioner(u, th, p),
                           rtol=1e-10, max_iter=3000)
    assert float(torch.norm(b - fem.make_A_operator_qp(u, th, p)(x_ref))) < 1e-9 * float(torch.norm(b))
    out = {}
    for ch in ("small", "large"):
        m = HipMesh(p, DEV, chunking=ch)
        assert bool(m.kind & _lib.SMALL) == (ch == "small")
        s = HipSystem(m, th, material_spec(p), p["f"], precision)
        s.moduli(u.to(DEV))
        y = s.apply(v.to(DEV).to(s.vdt))
        assert rel(y, ref) < tol
        assert rel(s.residual(u.to(DEV)), fem.global_residual(u, th, p)) < 1e-11
        assert rel(s.diag, fem.compute_jacobi_diagonal(u, th, p)) < (1e-11 if precision == "fp64" else 1e-6)
        x, k, st = s.pcg_solve(b.to(DEV).to(s.vdt), rtol=1e-10 if precision == "fp64" else 1e-6, max_iter=3000)
        assert st == _lib.PCG_CONVERGED
        assert rel(x, x_ref) < (1e-8 if precision == "fp64" else 1e-4)
        out[ch] = (y.double().cpu(), k)
    assert rel(out["small"][0], out["large"][0]) < tol
    assert abs(out["small"][1] - out["large"][1]) <= max(2, 0.02 * out["large"][1])


@pytest.mark.parametrize("et,n", [("quad4", 400), ("tri3", 283)])
def test_tail_balanced_and_plain_plans_agree_with_oracle(et, n):
    """Tail-balanced plans (whole rounds of 256-element chunks, the rest in half chunks) against the plain plan
    and the oracle: the product in FP64 and FP32, a PCG solve, and a batched product of two solves (more work
    items than the issue-ahead variant's slots: the gather-first kernel runs for tail plans at any launch size)."""
    import ctypes
    import numpy as np
    from difffea_amd import _lib
    from difffea_amd.device import HipMesh, HipSystem, _ptr, _stream
    from difffea_amd.materials import material_spec
    from difffea_amd.plan import MeshPlan
    p = fem.precompute_geometry(unit_square(et, n))
    u = stretch_state(p, 1.0 / n, seed=0)
    v = torch.randn(u.shape[0], dtype=torch.float64, generator=torch.Generator().manual_seed(1))
    th = torch.ones(p["conn"].shape[0], dtype=torch.float64)
    ref = fem.make_A_operator_qp(u, th, p)(v)
    A_ref = fem.make_A_operator_qp(u, th, p)
    b = v * p["m"]
    out = {}
    for balance in (True, False):
        plan = MeshPlan(p["nodes"].numpy(), p["conn"].numpy(), p["fixed"].numpy(), balance=balance,
                        cell_ij=p["cell_ij"].numpy() if "cell_ij" in p else None)
        assert plan.tail == balance
        m = HipMesh(p, DEV, plan=plan)
        assert bool(m.kind & _lib.TAIL) == balance
        for precision, tol in (("fp64", 1e-11), ("mixed", 1e-5)):
            s = HipSystem(m, th, material_spec(p), p["f"], precision)
            s.moduli(u.to(DEV))
            y = s.apply(v.to(DEV).to(s.vdt))
            assert rel(y, ref) < tol
            out[(balance, precision)] = y.double().cpu()
        s = HipSystem(m, th, material_spec(p), p["f"], "fp64")
        s.moduli(u.to(DEV))
        x, k, st = s.pcg_solve(b.to(DEV), rtol=1e-8, max_iter=5000)
        assert st == _lib.PCG_CONVERGED
        assert float(torch.norm(b - A_ref(x.cpu()))) <= 1.01e-8 * float(torch.norm(b))
        out[(balance, "pcg")] = (x.cpu(), k)
        if balance:
            # two solves of the tail plan in one launch (2 x 738 or 2 x 626 work items) give the single-solve bits
            s2 = HipSystem(m, th, material_spec(p), p["f"], "fp64")
            s2.moduli(u.to(DEV))
            w = torch.randn(u.shape[0], dtype=torch.float64, generator=torch.Generator().manual_seed(2)).to(DEV)
            y_alone = [s.apply(v.to(DEV)), s2.apply(w)]
            table = torch.cat([m.upload(s.desc), m.upload(s2.desc)])
            nc, nf = m.plan.n_chunks, m.n_finblk
            items = torch.tensor(np.r_[np.c_[np.zeros(nc), np.arange(nc)], np.c_[np.ones(nc), np.arange(nc)]]
                                 .astype(np.int32).reshape(-1), device=DEV)
            fin = torch.tensor(np.r_[np.c_[np.zeros(nf), np.arange(nf)], np.c_[np.ones(nf), np.arange(nf)]]
                               .astype(np.int32).reshape(-1), device=DEV)
            for sys_, vec in ((s, v.to(DEV)), (s2, w)):
                sys_.vec["p"].copy_(vec)
            assert 2 * nc > 768
            _lib.check(m.lib.dfea_apply(_ptr(table), _ptr(items), 2 * nc, _ptr(fin), 2 * nf, None, None, m.kind,
                                        _lib.FP64, _stream()), "dfea_apply")
            assert torch.equal(s.vec["w"], y_alone[0]) and torch.equal(s2.vec["w"], y_alone[1])
    assert rel(out[(True, "fp64")], out[(False, "fp64")]) < 1e-12
    (xa, ka), (xb, kb) = out[(True, "pcg")], out[(False, "pcg")]
    assert abs(ka - kb) <= max(2, 0.02 * kb) and rel(xa, xb) < 1e-6
