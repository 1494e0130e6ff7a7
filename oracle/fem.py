"""CPU oracle (FP64, PyTorch-CPU + torch.func) for the matrix-free Newton-Krylov hot path.

TEST INFRASTRUCTURE ONLY.  This module is imported by ``tests/``, by ``__graft_entry__.smoke()`` and by the
``cpu_baseline`` leg of ``bench.py`` -- as the checker, never as the thing measured or shipped.  The product
path (``difffea_amd``) never imports it and fails loudly when its HIP library is missing.

It restates, in its own code, the reference's algorithms step for step (file:line of the reference's
``src/difffea`` is cited on every function), including the reference's use of automatic differentiation
(stress = grad psi, moduli = hessian psi, tangent = jvp, transpose = vjp, diagonal = jvp on unit vectors).

How each function is pinned (tests/test_oracle_*.py):
  * geometry               -- affine/rectangle closed forms, shape-gradient finite differences, mesh area
  * library energies       -- stress-free reference, objectivity, plane-strain closed forms (Mooney-Rivlin
                              I2 = J^2 + F:F, SVK through tr C^2 = (tr C)^2 - 2 det C), limits (Demiray
                              gamma -> 0 and HGO k1 -> 0 are neo-Hookean), fibres slack in compression,
                              P and C against central differences, R_int = grad(total energy)
  * psi / P / C            -- closed forms P = mu F + (lam lnJ - mu) F^-T and
                              C = mu d_ik d_jl + lam G_ij G_kl + (mu - lam lnJ) G_il G_kj, linear-elastic
                              tensor at F = I, objectivity, stress-free reference state
  * global_residual        -- rigid-body invariance, R_int = grad(total energy), finite differences
  * qp / qp_ew / jvp / linearize A(v) -- mutual agreement < 1e-11, symmetry, finite differences of R
  * jacobi diagonal        -- diagonal of the dense operator
  * pcg                    -- dense solve
  * newton_krylov_solve    -- equilibrium, reactions balance the applied load, golden trajectory (G2)
  * adjoint                -- central finite differences; golden reference adjoint (G3)
  * block-Jacobi blocks    -- the dense operator's nodal 2x2 diagonal blocks; PCG with them = dense solve
  * fields / reactions     -- affine patches in closed form (Cauchy stress, energy, J, von Mises), element-by-
                              element loops for the nodal averages, reactions = -(applied load) at equilibrium
                              and = the fixed-dof sum of grad(total energy)
"""
import torch
from torch.func import grad, hessian, jvp, linearize, vjp, vmap

DT = torch.float64


# ---------------------------------------------------------------------------------------------------------
# geometry cache                                                    (operators.py:85-120, kinematics.py:78-104)
# ---------------------------------------------------------------------------------------------------------
def shape_fn_quad4(xi, eta):
    """Bilinear Q4, CCW corners (shape_functions_2d.py:77-97)."""
    return torch.stack([0.25 * (1 - xi) * (1 - eta), 0.25 * (1 + xi) * (1 - eta),
                        0.25 * (1 + xi) * (1 + eta), 0.25 * (1 - xi) * (1 + eta)])


def shape_fn_quad9(xi, eta):
    """Biquadratic Lagrange Q9: corners, mid-sides (bottom, right, top, left), centre (:132-169)."""
    a = [0.5 * xi * (xi - 1), 1 - xi ** 2, 0.5 * xi * (xi + 1)]
    b = [0.5 * eta * (eta - 1), 1 - eta ** 2, 0.5 * eta * (eta + 1)]
    return torch.stack([a[0] * b[0], a[2] * b[0], a[2] * b[2], a[0] * b[2],
                        a[1] * b[0], a[2] * b[1], a[1] * b[2], a[0] * b[1], a[1] * b[1]])


def shape_fn_quad8(xi, eta):
    """Serendipity Q8: corners with the (xi_a xi + eta_a eta - 1) factor, then mid-sides bottom, right, top,
    left (shape_functions_2d.py:100-129)."""
    corners = [(-1.0, -1.0), (1.0, -1.0), (1.0, 1.0), (-1.0, 1.0)]
    N = [0.25 * (1 + a * xi) * (1 + b * eta) * (a * xi + b * eta - 1) for a, b in corners]
    N += [0.5 * (1 - xi ** 2) * (1 - eta), 0.5 * (1 + xi) * (1 - eta ** 2),
          0.5 * (1 - xi ** 2) * (1 + eta), 0.5 * (1 - xi) * (1 - eta ** 2)]
    return torch.stack(N)


def shape_fn_tri3(xi, eta):
    """Linear triangle: the barycentric coordinates (1 - xi - eta, xi, eta) (shape_functions_2d.py:177-195)."""
    return torch.stack([1 - xi - eta, xi, eta])


def shape_fn_tri6(xi, eta):
    """Quadratic triangle: corners L(2L - 1), mid-sides of edges 1-2, 2-3, 3-1 as 4 L_i L_j
    (shape_functions_2d.py:198-230)."""
    L1, L2, L3 = 1 - xi - eta, xi, eta
    return torch.stack([L1 * (2 * L1 - 1), L2 * (2 * L2 - 1), L3 * (2 * L3 - 1), 4 * L1 * L2, 4 * L2 * L3,
                        4 * L3 * L1])


SHAPE_FN = {"quad4": shape_fn_quad4, "quad8": shape_fn_quad8, "quad9": shape_fn_quad9,
            "tri3": shape_fn_tri3, "tri6": shape_fn_tri6}


def gauss_rule_2d(n):
    """Tensor Gauss-Legendre rule, eta-major / xi-minor ordering (gauss_quadratures.py:30-94)."""
    import numpy as np
    p, w = np.polynomial.legendre.leggauss(n)
    p, w = torch.tensor(p, dtype=DT), torch.tensor(w, dtype=DT)
    GP = torch.stack([p.repeat(n), p.repeat_interleave(n)], dim=1)
    return GP, torch.outer(w, w).reshape(-1)


def gauss_rule_triangle(degree):
    """Rule on the unit triangle exact to total degree ``degree`` (gauss_quadratures.py:100-160): centroid (1),
    three interior points (2), the two-orbit 6-point rule with the reference's 15-digit constants (3-4), else
    the collapsed rule xi = s, eta = t (1 - s) from Gauss-Legendre points on [0, 1], s-major.  Weights sum to
    1/2."""
    import math
    import numpy as np
    if degree <= 1:
        pts, wts = [[1 / 3, 1 / 3]], [0.5]
    elif degree == 2:
        pts, wts = [[1 / 6, 1 / 6], [2 / 3, 1 / 6], [1 / 6, 2 / 3]], [1 / 6] * 3
    elif degree <= 4:
        pts, wts = [], []
        for a, w in ((0.445948490915965, 0.223381589678011), (0.091576213509771, 0.109951743655322)):
            pts += [[a, a], [1 - 2 * a, a], [a, 1 - 2 * a]]
            wts += [w / 2] * 3
    else:
        n = math.ceil((degree + 2) / 2)
        g, w = np.polynomial.legendre.leggauss(n)
        g, w = (g + 1) / 2, w / 2
        pts = [[gs, gt * (1 - gs)] for gs in g for gt in g]
        wts = [ws * wt * (1 - gs) for gs, ws in zip(g, w) for wt in w]
    return torch.tensor(pts, dtype=DT), torch.tensor(wts, dtype=DT)


def element_rule(element_type):
    """Default rule of an element type (elements.py:38-93): 2x2 quad4, 3x3 quad8/quad9, degree 1 tri3,
    degree 4 tri6."""
    return {"quad4": lambda: gauss_rule_2d(2), "quad8": lambda: gauss_rule_2d(3), "quad9": lambda: gauss_rule_2d(3),
            "tri3": lambda: gauss_rule_triangle(1), "tri6": lambda: gauss_rule_triangle(4)}[element_type]()


def reference_gradients(element_type, GP):
    """dN_a/dxi_j at the Gauss points by forward-mode AD of the shape functions (grad_shape_functions_2d.py)."""
    from torch.func import jacfwd
    fn = SHAPE_FN[element_type]
    return vmap(lambda x, e: torch.stack(jacfwd(fn, argnums=(0, 1))(x, e), dim=-1))(GP[:, 0], GP[:, 1])


def precompute_geometry(params):
    """dN_dX[e,q,a,j] = sum_k dN_a/dxi_k (J^-1)_{kj},  w_detJ[e,q] = w_q det J, J = dX/dxi (columns)."""
    et = params["element_type"]
    dN_dxi = reference_gradients(et, params["GP"].to(DT))                       # (n_q, nen, 2)
    Xe = params["nodes"][params["conn"]].to(DT)                                  # (n_el, nen, 2)
    J = torch.einsum("eai,qaj->eqij", Xe, dN_dxi)                                # J[i,j] = dX_i/dxi_j
    detJ = torch.linalg.det(J)
    invJ = torch.linalg.inv(J)
    dN_dX = torch.einsum("qaj,eqji->eqai", dN_dxi, invJ)
    return {**params, "dN_dX": dN_dX, "w_detJ": params["GW"].to(DT)[None, :] * detJ}


def get_geometry(params):
    if "dN_dX" in params:
        return params["dN_dX"], params["w_detJ"]
    g = precompute_geometry(params)
    return g["dN_dX"], g["w_detJ"]


# ---------------------------------------------------------------------------------------------------------
# material                                                          (materials.py:37-246)
# ---------------------------------------------------------------------------------------------------------
def psi_neo_hookean(F, p):
    """Plane-strain compressible neo-Hookean energy through the 3-D embedding F3 = [[F,0],[0,1]].

    psi = mu/2 (I1 - 3 - 2 ln J) + lam/2 (ln J)^2 with I1 = tr(F3^T F3) = F:F + 1 and J = det F (explicit
    2x2 determinant so that second derivatives stay finite at F = I).
    """
    mu, lam = p[0], p[1]
    I1 = torch.sum(F * F) + 1.0
    logJ = torch.log(F[0, 0] * F[1, 1] - F[0, 1] * F[1, 0])
    return 0.5 * mu * (I1 - 3.0 - 2.0 * logJ) + 0.5 * lam * logJ ** 2


def embed_plane_strain(F):
    """F (2,2) -> F3 = [[F, 0], [0, 1]] (`materials.py:37-48`)."""
    z, o = torch.zeros_like(F[0, 0]), torch.ones_like(F[0, 0])
    return torch.stack([torch.stack([F[0, 0], F[0, 1], z]), torch.stack([F[1, 0], F[1, 1], z]),
                        torch.stack([z, z, o])])


def _det3(F3):
    """det F3 of the block embedding, written out so second derivatives stay finite (`materials.py:51-54`)."""
    return (F3[0, 0] * F3[1, 1] - F3[0, 1] * F3[1, 0]) * F3[2, 2]


def psi3_neo_hookean(F3, p):
    """mu/2 (I1 - 3 - 2 ln J) + lam/2 (ln J)^2, p = [mu, lmbda] (`materials.py:65-73`)."""
    mu, lam = p[0], p[1]
    logJ = torch.log(_det3(F3))
    return 0.5 * mu * (torch.sum(F3 * F3) - 3.0 - 2.0 * logJ) + 0.5 * lam * logJ ** 2


def psi3_mooney_rivlin(F3, p):
    """c10 (J^-2/3 I1 - 3) + c01 (J^-4/3 I2 - 3) + kappa/2 (ln J)^2, I2 = (I1^2 - tr C^2)/2
    (`materials.py:76-93`)."""
    c10, c01, kappa = p[0], p[1], p[2]
    C = F3.T @ F3
    I1 = torch.trace(C)
    I2 = 0.5 * (I1 ** 2 - torch.trace(C @ C))
    J = _det3(F3)
    return c10 * (J ** (-2.0 / 3.0) * I1 - 3.0) + c01 * (J ** (-4.0 / 3.0) * I2 - 3.0) + 0.5 * kappa * torch.log(J) ** 2


def psi3_saint_venant(F3, p):
    """lam/2 (tr E)^2 + mu tr(E^2), E = (C - I)/2 (`materials.py:96-104`)."""
    mu, lam = p[0], p[1]
    E = 0.5 * (F3.T @ F3 - torch.eye(3, dtype=F3.dtype))
    return 0.5 * lam * torch.trace(E) ** 2 + mu * torch.sum(E * E)


def psi3_demiray(F3, p):
    """mu/(2 gamma) (exp(gamma (I1 - 3)) - 1) - mu ln J + lam/2 (ln J)^2 (`materials.py:107-119`)."""
    mu, gamma, lam = p[0], p[1], p[2]
    I1 = torch.sum(F3 * F3)
    logJ = torch.log(_det3(F3))
    return mu / (2.0 * gamma) * (torch.exp(gamma * (I1 - 3.0)) - 1.0) - mu * logJ + 0.5 * lam * logJ ** 2


def psi3_fiber_reinforced(F3, p):
    """Neo-Hookean matrix + fibre families a = (cos b, +-sin b, 0): k1/(2 k2) (exp(k2 <E>^2) - 1),
    E = d I1 + (1 - 3d) a.C a - 1, <.> = max(., 0) (`materials.py:122-162`)."""
    k1, k2, beta, d, second = p[2], p[3], torch.deg2rad(p[4]), p[5], p[6]
    C = F3.T @ F3
    I1 = torch.trace(C)
    z = torch.zeros_like(beta)

    def family(a):
        E = d * I1 + (1.0 - 3.0 * d) * (a @ C @ a) - 1.0
        return k1 / (2.0 * k2) * (torch.exp(k2 * torch.relu(E) ** 2) - 1.0)

    a1 = torch.stack([torch.cos(beta), torch.sin(beta), z])
    a2 = torch.stack([torch.cos(beta), -torch.sin(beta), z])
    return psi3_neo_hookean(F3, p[:2]) + family(a1) + second * family(a2)


PSI3 = {"neo_hookean": psi3_neo_hookean, "mooney_rivlin": psi3_mooney_rivlin, "saint_venant": psi3_saint_venant,
        "demiray": psi3_demiray, "fiber_reinforced": psi3_fiber_reinforced}


def material_psi(params):
    """The plane-strain energy psi(F2, p) of the problem's material (`materials.py:199-205`); the neo-Hookean
    one in its 2-D form above."""
    name = params.get("material", "neo_hookean")
    if name == "neo_hookean":
        return psi_neo_hookean
    psi3 = PSI3[name]
    return lambda F, p: psi3(embed_plane_strain(F), p)


def element_material_parameters(theta, params):
    """(n_el, 2): nominal [mu, lam] with the scalable entries (both, for neo-Hookean) times theta_e."""
    p0 = params["mat_p"].to(theta.dtype)
    sc = params["mat_scalable"]
    factor = torch.where(sc[None, :], theta[:, None], torch.ones_like(theta)[:, None])
    return p0[None, :] * factor


# ---------------------------------------------------------------------------------------------------------
# residual                                                          (operators.py:131-220)
# ---------------------------------------------------------------------------------------------------------
def elem_residual_precomputed(ue, dN_dX, w_detJ, p, psi=psi_neo_hookean):
    """R^e[a,i] = sum_q w_q detJ_q P_ij(F_q) dN_a/dX_j with P = grad psi (automatic differentiation)."""
    F = torch.eye(2, dtype=ue.dtype) + torch.einsum("ai,qaj->qij", ue, dN_dX)
    P = vmap(grad(psi), in_dims=(0, None))(F, p)
    return torch.einsum("q,qij,qaj->ai", w_detJ, P, dN_dX)


def global_residual(u, theta, params, load_factor=1.0):
    """R = m * (scatter_e R^e(u) - lf f)."""
    conn = params["conn"]
    N = params["nodes"].shape[0]
    ue = u.reshape(N, 2)[conn]
    dN_dX, w_detJ = get_geometry(params)
    pe = element_material_parameters(theta, params)
    psi = material_psi(params)
    Re = vmap(lambda a, b, c, d: elem_residual_precomputed(a, b, c, d, psi))(ue, dN_dX, w_detJ, pe)
    Rint = torch.zeros(N, 2, dtype=u.dtype).index_add_(0, conn.reshape(-1), Re.reshape(-1, 2))
    return params["m"] * (Rint.reshape(-1) - load_factor * params["f"])


# ---------------------------------------------------------------------------------------------------------
# tangent operators                                                 (operators.py:231-435)
# ---------------------------------------------------------------------------------------------------------
def make_A_operator(u, theta, params):
    """'jvp': A(v) = m J(u)(m v) + (1-m) v with J v by forward-mode AD of the residual."""
    m = params["m"]

    def A(v):
        _, Jv = jvp(lambda w: global_residual(w, theta, params), (u,), (m * v,))
        return m * Jv + (1.0 - m) * v
    return A


def make_A_operator_linearize(u, theta, params):
    """'linearize' (operators.py:299-313): the residual is linearised once at u (torch.func.linearize keeps the
    primal evaluation), and each product pushes m v through the stored linearisation; the same operator as
    'jvp'."""
    m = params["m"]
    _, lin = linearize(lambda w: global_residual(w, theta, params), u)

    def A(v):
        return m * lin(m * v) + (1.0 - m) * v
    return A


def qp_setup(u, theta, params):
    """C[e,q,2i+j,2k+l] = w_detJ[e,q] * d2psi/dF_ij dF_kl (hessian by AD), and dN_dX."""
    conn = params["conn"]
    N = params["nodes"].shape[0]
    n_el = conn.shape[0]
    dN_dX, w_detJ = get_geometry(params)
    n_q = dN_dX.shape[1]
    pe = element_material_parameters(theta, params)
    ue = u.reshape(N, 2)[conn]
    F = torch.eye(2, dtype=u.dtype) + torch.einsum("eai,eqaj->eqij", ue, dN_dX)
    H = vmap(vmap(hessian(material_psi(params)), in_dims=(0, None)), in_dims=(0, 0))(F, pe)
    return (H * w_detJ[:, :, None, None, None, None]).reshape(n_el, n_q, 4, 4), dN_dX


def make_A_operator_qp(u, theta, params, C=None):
    """'qp': gather -> einsum -> batched 4x4 matmul -> einsum -> index_add_ -> identity pivot.

    This op-for-op restatement is also the timed CPU baseline of the tangent product.
    """
    conn, m = params["conn"], params["m"]
    N = params["nodes"].shape[0]
    n_el = conn.shape[0]
    if C is None:
        C, dN_dX = qp_setup(u, theta, params)
    else:
        dN_dX, _ = get_geometry(params)
    n_q = dN_dX.shape[1]
    flat = conn.reshape(-1)

    def A(v):
        ve = (m * v).reshape(N, 2)[conn]
        g = torch.einsum("eai,eqaj->eqij", ve, dN_dX).reshape(n_el, n_q, 4, 1)
        s = torch.matmul(C, g).reshape(n_el, n_q, 2, 2)
        fe = torch.einsum("eqij,eqaj->eai", s, dN_dX)
        Jv = torch.zeros(N, 2, dtype=v.dtype).index_add_(0, flat, fe.reshape(-1, 2))
        return m * Jv.reshape(-1) + (1.0 - m) * v
    return A


def make_A_operator_qp_ew(u, theta, params):
    """'qp_ew': the same product with broadcast multiply + sum instead of matmul."""
    conn, m = params["conn"], params["m"]
    N = params["nodes"].shape[0]
    n_el = conn.shape[0]
    C, dN_dX = qp_setup(u, theta, params)
    n_q = dN_dX.shape[1]
    flat = conn.reshape(-1)

    def A(v):
        ve = (m * v).reshape(N, 2)[conn]
        g = (ve[:, None, :, :, None] * dN_dX[:, :, :, None, :]).sum(2)
        s = (C * g.reshape(n_el, n_q, 1, 4)).sum(-1).reshape(n_el, n_q, 2, 2)
        fe = (s[:, :, None, :, :] * dN_dX[:, :, :, None, :]).sum(dim=(1, 4))
        Jv = torch.zeros(N, 2, dtype=v.dtype).index_add_(0, flat, fe.reshape(-1, 2))
        return m * Jv.reshape(-1) + (1.0 - m) * v
    return A


ORACLE_TANGENTS = ("jvp", "linearize", "qp", "qp_ew")


def make_tangent_operator(kind, u, theta, params):
    if kind in ("auto", "qp"):
        return make_A_operator_qp(u, theta, params)
    if kind == "qp_ew":
        return make_A_operator_qp_ew(u, theta, params)
    if kind == "jvp":
        return make_A_operator(u, theta, params)
    if kind == "linearize":
        return make_A_operator_linearize(u, theta, params)
    raise ValueError(f"Unknown tangent operator '{kind}'. Choose one of {ORACLE_TANGENTS}.")


def make_AT_operator(u, theta, params):
    """AT(v) = m J^T(m v) + (1-m) v by reverse-mode AD (operators.py:441-471)."""
    m = params["m"]
    _, f_vjp = vjp(lambda w: global_residual(w, theta, params), u)

    def AT(v):
        (Jt,) = f_vjp(m * v)
        return m * Jt + (1.0 - m) * v
    return AT


# ---------------------------------------------------------------------------------------------------------
# Jacobi diagonal                                                   (operators.py:477-570)
# ---------------------------------------------------------------------------------------------------------
def compute_jacobi_diagonal(u, theta, params):
    """Per element: e_k . (J^e e_k) for the 2 nen local unit vectors (jvp), scatter, clamp / pivot."""
    conn = params["conn"]
    N = params["nodes"].shape[0]
    n_el, nen = conn.shape
    ue = u.reshape(N, 2)[conn]
    pe = element_material_parameters(theta, params)
    dN_dX, w_detJ = get_geometry(params)
    basis = torch.eye(2 * nen, dtype=u.dtype).reshape(-1, nen, 2)

    psi = material_psi(params)

    def entry(u_e, dN_e, w_e, p_e, e_vec):
        _, Je = jvp(lambda w: elem_residual_precomputed(w, dN_e, w_e, p_e, psi), (u_e,), (e_vec,))
        return (Je * e_vec).sum()

    def one(u_e, dN_e, w_e, p_e):
        return vmap(entry, in_dims=(None, None, None, None, 0))(u_e, dN_e, w_e, p_e, basis)

    de = vmap(one)(ue, dN_dX, w_detJ, pe).reshape(n_el, nen, 2)
    d = torch.zeros(N, 2, dtype=u.dtype).index_add_(0, conn.reshape(-1), de.reshape(-1, 2)).reshape(-1)
    return torch.where(params["m"] > 0.0, torch.clamp(d, min=1e-6), torch.ones_like(d))


def make_jacobi_preconditioner(u, theta, params):
    diag = compute_jacobi_diagonal(u, theta, params)
    return lambda r: r / diag


# ---------------------------------------------------------------------------------------------------------
# Krylov + Newton                                                   (newton_krylov.py:28-264)
# ---------------------------------------------------------------------------------------------------------
def element_tangent_blocks(u, theta, params):
    """The 2x2 diagonal blocks of every element tangent, (n_el, nen, 2, 2) [e, a, i, j] = dR^e[a,i]/du[a,j]
    (`operators.py:588-615`): the element Jacobian column by column (one jvp per local unit vector), then its
    node-diagonal blocks."""
    conn = params["conn"]
    N = params["nodes"].shape[0]
    nen = conn.shape[1]
    ue = u.reshape(N, 2)[conn]
    pe = element_material_parameters(theta, params)
    dN_dX, w_detJ = get_geometry(params)
    units = torch.eye(2 * nen, dtype=u.dtype).reshape(-1, nen, 2)

    psi = material_psi(params)

    def jacobian(u_e, dN_e, w_e, p_e):
        res = lambda w: elem_residual_precomputed(w, dN_e, w_e, p_e, psi)
        cols = vmap(lambda e: jvp(res, (u_e,), (e,))[1])(units)      # [(b, j), a, i]
        cols = cols.reshape(nen, 2, nen, 2)                           # [b, j, a, i]
        return torch.diagonal(cols, dim1=0, dim2=2).permute(2, 1, 0)  # [a, i, j], a == b

    return vmap(jacobian)(ue, dN_dX, w_detJ, pe)


def block_jacobi_blocks(u, theta, params):
    """Nodal 2x2 blocks of A: scatter of the element blocks, identity rows / columns on fixed dofs, plus the
    1e-12 I guard (`operators.py:618-638`); returns (N, 2, 2)."""
    conn, m = params["conn"], params["m"]
    N = params["nodes"].shape[0]
    blocks = element_tangent_blocks(u, theta, params)
    B = torch.zeros(N, 2, 2, dtype=u.dtype).index_add_(0, conn.reshape(-1), blocks.reshape(-1, 2, 2))
    mn = m.reshape(N, 2)
    B = B * (mn[:, :, None] * mn[:, None, :]) + torch.diag_embed(1.0 - mn)
    return B + 1e-12 * torch.eye(2, dtype=u.dtype)


def make_block_jacobi_preconditioner(u, theta, params):
    """r -> B_n^-1 r_n per node (`operators.py:618-638`)."""
    N = params["nodes"].shape[0]
    Binv = torch.linalg.inv(block_jacobi_blocks(u, theta, params))
    return lambda r: torch.einsum("nij,nj->ni", Binv, r.reshape(N, 2)).reshape(-1)


PRECONDITIONERS = {"jacobi": make_jacobi_preconditioner, "block_jacobi": make_block_jacobi_preconditioner}


def pcg(A_fn, b, Minv_fn, rtol=1e-5, max_iter=500):
    """Textbook Jacobi-PCG with x0 = 0; curvature guard; stop when ||r|| <= rtol ||b||."""
    x = torch.zeros_like(b)
    r = b.clone()
    z = Minv_fn(r)
    p = z.clone()
    rz = torch.dot(r, z)
    nb = torch.norm(b)
    if nb == 0.0:
        return x, 0
    for k in range(max_iter):
        Ap = A_fn(p)
        curv = torch.dot(p, Ap)
        if curv <= 0.0:
            return (z if k == 0 else x), k
        alpha = rz / curv
        x = x + alpha * p
        r = r - alpha * Ap
        if torch.norm(r) <= rtol * nb:
            return x, k + 1
        z = Minv_fn(r)
        rz_new = torch.dot(r, z)
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x, max_iter


def armijo_line_search(u, du, rn, theta, params, load_factor, c1=1e-4, min_alpha=2.0 ** -5):
    alpha = 1.0
    while alpha >= min_alpha:
        uc = u + alpha * du
        Rc = global_residual(uc, theta, params, load_factor)
        rc = torch.norm(Rc)
        if rc < (1.0 - c1 * alpha) * rn:
            return uc, Rc, rc, alpha, True
        alpha *= 0.5
    return u, None, rn, alpha, False


def newton_at_load(u0, theta, params, lf, tol, max_newton, step_id, tangent="qp",
                   cg_rtol=None, cg_max_iter=500, callback=None, preconditioner="jacobi"):
    u = torch.where(params["fixed"], lf * params["u_D"], u0)
    R = global_residual(u, theta, params, lf)
    r0 = torch.norm(R)
    rn = r0
    records = []
    for it in range(max_newton + 1):
        if rn < tol:
            return u, True, records
        if it == max_newton:
            break
        eta = min(0.5, float((rn / (r0 + 1e-14)) ** 0.5)) if cg_rtol is None else cg_rtol
        A = make_tangent_operator(tangent, u, theta, params)
        Minv = PRECONDITIONERS[preconditioner](u, theta, params)
        du, k = pcg(A, -R, Minv, rtol=eta, max_iter=cg_max_iter)
        u, Rn, rn, alpha, ok = armijo_line_search(u, du, rn, theta, params, lf)
        rec = dict(step=step_id, load_factor=lf, iteration=it, residual=float(rn), cg_iters=int(k), alpha=alpha)
        records.append(rec)
        if callback is not None:
            callback(rec)
        if not ok:
            return u, False, records
        R = Rn
    return u, False, records


@torch.no_grad()
def newton_krylov_solve(params, theta, n_load_steps=5, tol=1e-8, max_newton=25, max_cutbacks=8,
                        tangent="qp", cg_rtol=None, cg_max_iter=500, callback=None, preconditioner="jacobi"):
    """Adaptive load continuation around ``newton_at_load``; returns (u, info) like the reference."""
    n = 2 * params["nodes"].shape[0]
    u = torch.zeros(n, dtype=DT)
    info = {"converged": False, "load_factor": 0.0, "history": [], "total_newton": 0, "total_cg": 0,
            "n_cutbacks": 0, "snapshots": [(0.0, u.clone())]}
    d_nom = 1.0 / n_load_steps
    d = d_nom
    lf = 0.0
    step, in_row = 0, 0
    while lf < 1.0 - 1e-12:
        lf_try = min(1.0, lf + d)
        step += 1
        u_try, ok, recs = newton_at_load(u, theta, params, lf_try, tol, max_newton, step, tangent,
                                         cg_rtol, cg_max_iter, callback, preconditioner)
        info["history"] += recs
        info["total_newton"] += len(recs)
        info["total_cg"] += sum(r["cg_iters"] for r in recs)
        if ok:
            u, lf = u_try, lf_try
            info["snapshots"].append((lf, u.clone()))
            in_row = 0
            d = min(d_nom, 1.5 * d)
        else:
            in_row += 1
            info["n_cutbacks"] += 1
            if in_row > max_cutbacks:
                info["load_factor"] = lf
                return u, info
            d *= 0.5
    info["converged"] = True
    info["load_factor"] = 1.0
    return u, info


# ---------------------------------------------------------------------------------------------------------
# adjoint sensitivities                                             (adjoint.py:31-106, sensitivity.py:7-38)
# ---------------------------------------------------------------------------------------------------------
def solve_adjoint(u_star, theta, params, dL_du, rtol=1e-6, max_iter=500):
    AT = make_AT_operator(u_star, theta, params)
    Minv = make_jacobi_preconditioner(u_star, theta, params)
    return pcg(AT, params["m"] * dL_du, Minv, rtol=rtol, max_iter=max_iter)


def compute_implicit_vjp(u_star, theta, params, lam):
    _, f_vjp = vjp(lambda th: global_residual(u_star, th, params, 1.0), theta)
    (g,) = f_vjp(-lam)
    return g


def compute_sensitivity(params, theta, loss_fn, n_load_steps=5, tol=1e-8, max_newton=25,
                        adjoint_rtol=1e-10, adjoint_max_iter=500):
    theta = theta.detach()
    u, _ = newton_krylov_solve(params, theta, n_load_steps, tol, max_newton)
    dL_du = grad(loss_fn)(u)
    lam, _ = solve_adjoint(u, theta, params, dL_du, rtol=adjoint_rtol, max_iter=adjoint_max_iter)
    return u, compute_implicit_vjp(u, theta, params, lam)


# ---------------------------------------------------------------------------------------------------------
# post-processing                                                   (postprocess.py:24-137, materials.py:37-73)
# ---------------------------------------------------------------------------------------------------------
def gauss_point_fields(F, p, psi3=None):
    """[J, sxx, syy, sxy, szz, psi] at one Gauss point: P3 = grad psi3 (automatic differentiation),
    sigma = P3 F3^T / J (`postprocess.py:24-48`)."""
    psi3 = psi3 or psi3_neo_hookean
    F3 = embed_plane_strain(F)
    P3 = grad(psi3)(F3, p)
    J = F[0, 0] * F[1, 1] - F[0, 1] * F[1, 0]
    sig = P3 @ F3.T / J
    return torch.stack([J, sig[0, 0], sig[1, 1], sig[0, 1], sig[2, 2], psi3(F3, p)])


def compute_fields(u, theta, params):
    """Element means (weights w_q det J_q), von Mises of the mean stress, nodal averages over the elements
    at each node, total strain energy and min J (`postprocess.py:51-118`)."""
    conn = params["conn"]
    N = params["nodes"].shape[0]
    un = u.reshape(N, 2)
    dN_dX, w_detJ = get_geometry(params)
    F = torch.eye(2, dtype=u.dtype) + torch.einsum("eai,eqaj->eqij", un[conn], dN_dX)
    pe = element_material_parameters(theta, params)
    psi3 = PSI3[params.get("material", "neo_hookean")]
    S = vmap(vmap(lambda F_, p_: gauss_point_fields(F_, p_, psi3), in_dims=(0, None)))(F, pe)   # (n_el, n_q, 6)
    area = w_detJ.sum(1)
    J, sxx, syy, sxy, szz, psi = (torch.einsum("eq,eq->e", w_detJ, S[:, :, k]) / area for k in range(6))
    mises = torch.sqrt(0.5 * ((sxx - syy) ** 2 + (syy - szz) ** 2 + (szz - sxx) ** 2) + 3.0 * sxy ** 2)
    out = {"area": area, "J": J, "sxx": sxx, "syy": syy, "sxy": sxy, "mises": mises, "energy_density": psi,
           "ux": un[:, 0], "uy": un[:, 1], "u_mag": torch.linalg.norm(un, dim=1),
           "strain_energy": torch.sum(psi * area), "min_J": J.min()}
    for k in ("ux", "uy", "u_mag"):
        out[k + "_n"] = out[k]
    flat = conn.reshape(-1)
    count = torch.zeros(N, dtype=u.dtype).index_add_(0, flat, torch.ones(flat.numel(), dtype=u.dtype))
    for k in ("J", "sxx", "syy", "sxy", "mises", "energy_density"):
        s = torch.zeros(N, dtype=u.dtype).index_add_(0, flat, out[k][:, None].expand(-1, conn.shape[1]).reshape(-1))
        out[k + "_n"] = s / count.clamp(min=1.0)
    return out


def support_reactions(u, theta, params):
    """(Rx, Ry): the internal force (all-ones mask, no load) summed over the fixed dofs
    (`postprocess.py:121-137`)."""
    q = dict(params, m=torch.ones_like(params["m"]), f=torch.zeros_like(params["f"]))
    R = global_residual(u, theta, q, 1.0).reshape(-1, 2)
    fixed = params["fixed"].reshape(-1, 2)
    return torch.where(fixed, R, torch.zeros_like(R)).sum(0)
