"""Element tables for the two element types on the hot path: Q4 (2x2 Gauss) and Q9 (3x3 Gauss).

Host-side set-up only (CPU, float64).  Conventions follow the reference element library:

* node order Q4 = CCW corners (`src/difffea/shape_functions_2d.py:77-97`),
  Q9 = corners, then mid-sides bottom/right/top/left, then centre (`shape_functions_2d.py:132-169`);
* tensor Gauss rule ordered eta-major, xi-minor (`src/difffea/gauss_quadratures.py:64-94`);
* element edges as (start, end[, mid]) local ids (`src/difffea/elements.py:39-49`).

Shape-function gradients are written in closed form here (the reference obtains them with ``jacfwd``,
`src/difffea/grad_shape_functions_2d.py:39-61`); the tests compare them with finite differences.
"""
import numpy as np

ELEMENT_TYPES = ("quad4", "quad9")

# local edges: (start vertex, end vertex[, mid-side node])
EDGES = {
    "quad4": [(0, 1), (1, 2), (2, 3), (3, 0)],
    "quad9": [(0, 1, 4), (1, 2, 5), (2, 3, 6), (3, 0, 7)],
}
NEN = {"quad4": 4, "quad9": 9}
GAUSS_1D_POINTS = {"quad4": 2, "quad9": 3}

# reference coordinates of the nodes, (nen, 2)
REFERENCE_NODES = {
    "quad4": np.array([(-1, -1), (1, -1), (1, 1), (-1, 1)], dtype=np.float64),
    "quad9": np.array([(-1, -1), (1, -1), (1, 1), (-1, 1), (0, -1), (1, 0), (0, 1), (-1, 0), (0, 0)],
                      dtype=np.float64),
}


def element_type_of(nen):
    """Element type from the number of nodes per element (4 -> quad4, 9 -> quad9)."""
    for name, n in NEN.items():
        if n == nen:
            return name
    raise ValueError(f"no hot-path element type has {nen} nodes; supported: {ELEMENT_TYPES}")


def gauss_1d(n):
    """Gauss-Legendre points and weights on [-1, 1] (Golub-Welsch through numpy)."""
    pts, wts = np.polynomial.legendre.leggauss(n)
    return pts.astype(np.float64), wts.astype(np.float64)


def gauss_2d(n):
    """Tensor rule on [-1,1]^2: point k = (xi_i, eta_j) with k = j*n + i (eta-major, xi-minor).

    Returns GP (n*n, 2) [xi, eta] and GW (n*n,).
    """
    p, w = gauss_1d(n)
    xi = np.tile(p, n)
    eta = np.repeat(p, n)
    GP = np.stack([xi, eta], axis=1)
    GW = np.outer(w, w).reshape(-1)          # w[eta_j] * w[xi_i], same flattening as the points
    return GP, GW


def element_quadrature(element_type):
    """Default Gauss rule of an element type: 2x2 for quad4, 3x3 for quad9."""
    return gauss_2d(GAUSS_1D_POINTS[element_type])


def _lagrange1_1d(s):
    return np.stack([0.5 * (1.0 - s), 0.5 * (1.0 + s)]), np.stack([-0.5 * np.ones_like(s), 0.5 * np.ones_like(s)])


def _lagrange2_1d(s):
    """Quadratic Lagrange polynomials at s = -1, 0, +1 and their derivatives."""
    L = np.stack([0.5 * s * (s - 1.0), 1.0 - s * s, 0.5 * s * (s + 1.0)])
    dL = np.stack([s - 0.5, -2.0 * s, s + 0.5])
    return L, dL


# Q9 node -> (index of the xi polynomial, index of the eta polynomial); polynomial 0 at -1, 1 at 0, 2 at +1
_Q9_TENSOR = [(0, 0), (2, 0), (2, 2), (0, 2), (1, 0), (2, 1), (1, 2), (0, 1), (1, 1)]
_Q4_TENSOR = [(0, 0), (1, 0), (1, 1), (0, 1)]


def shape_values(element_type, xi, eta):
    """N_a(xi, eta) for arrays xi, eta of equal shape -> (..., nen)."""
    xi, eta = np.asarray(xi, np.float64), np.asarray(eta, np.float64)
    if element_type == "quad4":
        Lx, _ = _lagrange1_1d(xi)
        Le, _ = _lagrange1_1d(eta)
        tab = _Q4_TENSOR
    elif element_type == "quad9":
        Lx, _ = _lagrange2_1d(xi)
        Le, _ = _lagrange2_1d(eta)
        tab = _Q9_TENSOR
    else:
        raise ValueError(f"Unknown element type '{element_type}'. Choose one of {ELEMENT_TYPES}.")
    return np.stack([Lx[i] * Le[j] for i, j in tab], axis=-1)


def shape_gradients(element_type, xi, eta):
    """dN_a/d(xi, eta) for arrays xi, eta -> (..., nen, 2)."""
    xi, eta = np.asarray(xi, np.float64), np.asarray(eta, np.float64)
    if element_type == "quad4":
        Lx, dLx = _lagrange1_1d(xi)
        Le, dLe = _lagrange1_1d(eta)
        tab = _Q4_TENSOR
    elif element_type == "quad9":
        Lx, dLx = _lagrange2_1d(xi)
        Le, dLe = _lagrange2_1d(eta)
        tab = _Q9_TENSOR
    else:
        raise ValueError(f"Unknown element type '{element_type}'. Choose one of {ELEMENT_TYPES}.")
    d_xi = np.stack([dLx[i] * Le[j] for i, j in tab], axis=-1)
    d_eta = np.stack([Lx[i] * dLe[j] for i, j in tab], axis=-1)
    return np.stack([d_xi, d_eta], axis=-1)


def reference_gradient_table(element_type, GP=None):
    """dN_a/dxi_j at every Gauss point: (n_q, nen, 2) float64 (the kernels' small constant table)."""
    if GP is None:
        GP, _ = element_quadrature(element_type)
    GP = np.asarray(GP, np.float64)
    return shape_gradients(element_type, GP[:, 0], GP[:, 1])


def edge_shape(order, s):
    """1-D edge shape functions (end, end[, mid]) and derivatives at s: (..., k)."""
    s = np.asarray(s, np.float64)
    if order == 1:
        N, dN = _lagrange1_1d(s)
        return np.moveaxis(N, 0, -1), np.moveaxis(dN, 0, -1)
    L, dL = _lagrange2_1d(s)
    # edge ordering: start (s=-1), end (s=+1), mid (s=0)
    return np.moveaxis(L[[0, 2, 1]], 0, -1), np.moveaxis(dL[[0, 2, 1]], 0, -1)


def element_order(element_type):
    return 1 if element_type == "quad4" else 2
