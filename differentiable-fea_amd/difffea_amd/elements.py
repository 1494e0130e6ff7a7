"""Element tables of the hot path: tri3, quad4, tri6, quad8 and quad9 (host-side set-up, CPU float64).

Conventions follow the reference element library (`src/difffea/elements.py:38-57`):

* node orders: Q4 = CCW corners (`src/difffea/shape_functions_2d.py:77-97`); Q8 = corners, then mid-sides
  bottom/right/top/left (`:100-129`); Q9 = Q8 + centre (`:132-169`); Tri3 = (0,0), (1,0), (0,1)
  (`:177-195`); Tri6 = Tri3 corners, then mid-sides of edges 1-2, 2-3, 3-1 (`:198-230`);
* quadrature: tensor Gauss rule ordered eta-major, xi-minor for quadrilaterals (2x2 for Q4, 3x3 for Q8 and
  Q9; `src/difffea/gauss_quadratures.py:64-94`), symmetric triangle rules (centroid for Tri3, the 6-point
  degree-4 rule for Tri6; `gauss_quadratures.py:100-160`), weights summing to the reference-element area;
* element edges as (start, end[, mid]) local ids, CCW around the element (`src/difffea/elements.py:39-57`).

Shape-function gradients are written in closed form here (the reference obtains them with ``jacfwd``,
`src/difffea/grad_shape_functions_2d.py:39-61`); the tests compare them with finite differences.
"""
import numpy as np

ELEMENT_TYPES = ("quad4", "quad8", "quad9", "tri3", "tri6")

# local edges: (start vertex, end vertex[, mid-side node])
EDGES = {
    "quad4": [(0, 1), (1, 2), (2, 3), (3, 0)],
    "quad8": [(0, 1, 4), (1, 2, 5), (2, 3, 6), (3, 0, 7)],
    "quad9": [(0, 1, 4), (1, 2, 5), (2, 3, 6), (3, 0, 7)],
    "tri3": [(0, 1), (1, 2), (2, 0)],
    "tri6": [(0, 1, 3), (1, 2, 4), (2, 0, 5)],
}
NEN = {"quad4": 4, "quad8": 8, "quad9": 9, "tri3": 3, "tri6": 6}
FAMILY = {"quad4": "quad", "quad8": "quad", "quad9": "quad", "tri3": "tri", "tri6": "tri"}
ORDER = {"quad4": 1, "quad8": 2, "quad9": 2, "tri3": 1, "tri6": 2}
# default rule: ("tensor", points per direction) or ("triangle", polynomial degree)
GAUSS = {"quad4": ("tensor", 2), "quad8": ("tensor", 3), "quad9": ("tensor", 3),
         "tri3": ("triangle", 1), "tri6": ("triangle", 4)}

# reference coordinates of the nodes, (nen, 2)
REFERENCE_NODES = {
    "quad4": np.array([(-1, -1), (1, -1), (1, 1), (-1, 1)], dtype=np.float64),
    "quad8": np.array([(-1, -1), (1, -1), (1, 1), (-1, 1), (0, -1), (1, 0), (0, 1), (-1, 0)], dtype=np.float64),
    "quad9": np.array([(-1, -1), (1, -1), (1, 1), (-1, 1), (0, -1), (1, 0), (0, 1), (-1, 0), (0, 0)],
                      dtype=np.float64),
    "tri3": np.array([(0, 0), (1, 0), (0, 1)], dtype=np.float64),
    "tri6": np.array([(0, 0), (1, 0), (0, 1), (0.5, 0), (0.5, 0.5), (0, 0.5)], dtype=np.float64),
}


def _check(element_type):
    if element_type not in NEN:
        raise ValueError(f"Unknown element type '{element_type}'. Choose one of {list(ELEMENT_TYPES)}.")


def element_names():
    """Names of the element types (`src/difffea/elements.py:60-62`)."""
    return list(ELEMENT_TYPES)


def get_element(name):
    """The library entry of an element type: n_nodes, family, order, n_corners, edges, gauss
    (`src/difffea/elements.py:65-70`); unknown names raise ``ValueError``."""
    _check(name)
    return dict(n_nodes=NEN[name], family=FAMILY[name], order=ORDER[name],
                n_corners=3 if FAMILY[name] == "tri" else 4, edges=list(EDGES[name]), gauss=GAUSS[name])


def reference_node_coordinates(name):
    """(xi, eta) of the nodes on the reference element, (nen, 2) (`src/difffea/elements.py:103-118`)."""
    _check(name)
    return REFERENCE_NODES[name].copy()


def element_type_of(nen):
    """Element type from the number of nodes per element (the node counts of the five types differ)."""
    for name, n in NEN.items():
        if n == nen:
            return name
    raise ValueError(f"no element type has {nen} nodes; supported: {ELEMENT_TYPES}")


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


# Published symmetric degree-4 rule on the triangle (Dunavant 1985), two orbits (a, a, 1-2a) of three points;
# the weights add to 1 and are halved for the reference triangle's area.  The constants are the 15-digit
# values the reference uses (`src/difffea/gauss_quadratures.py:136-145`), so the Gauss points agree.
_DUNAVANT4 = ((0.445948490915965, 0.223381589678011), (0.091576213509771, 0.109951743655322))


def gauss_triangle(degree):
    """Rule on the reference triangle {xi, eta >= 0, xi + eta <= 1} exact for total degree <= ``degree``.

    degree 1: centroid; 2: three interior points (1/6, 1/6), (2/3, 1/6), (1/6, 2/3); 3-4: the 6-point
    symmetric rule, each orbit listed as (a, a), (1-2a, a), (a, 1-2a); higher: the collapsed (Duffy) rule
    xi = s, eta = t (1 - s) from an n x n Gauss rule on [0,1]^2 with n = ceil((degree + 2) / 2), s-major.
    Returns GP (n_q, 2), GW (n_q,), weights summing to 1/2.
    """
    if degree <= 1:
        return np.array([[1.0 / 3.0, 1.0 / 3.0]]), np.array([0.5])
    if degree == 2:
        GP = np.array([[1.0 / 6.0, 1.0 / 6.0], [2.0 / 3.0, 1.0 / 6.0], [1.0 / 6.0, 2.0 / 3.0]])
        return GP, np.full(3, 1.0 / 6.0)
    if degree <= 4:
        pts, wts = [], []
        for a, w in _DUNAVANT4:
            b = 1.0 - 2.0 * a
            pts += [[a, a], [b, a], [a, b]]
            wts += [0.5 * w] * 3
        return np.array(pts), np.array(wts)
    n = int(np.ceil((degree + 2) / 2))
    g, w = gauss_1d(n)
    g, w = 0.5 * (g + 1.0), 0.5 * w
    s = np.repeat(g, n)
    t = np.tile(g, n)
    GP = np.stack([s, t * (1.0 - s)], axis=1)
    GW = np.repeat(w, n) * np.tile(w, n) * (1.0 - s)
    return GP, GW


def element_quadrature(element_type):
    """Default rule of an element type: 2x2 (quad4), 3x3 (quad8, quad9), centroid (tri3), degree 4 (tri6)."""
    _check(element_type)
    kind, n = GAUSS[element_type]
    return gauss_2d(n) if kind == "tensor" else gauss_triangle(n)


def _lagrange1_1d(s):
    return np.stack([0.5 * (1.0 - s), 0.5 * (1.0 + s)]), np.stack([-0.5 * np.ones_like(s), 0.5 * np.ones_like(s)])


def _lagrange2_1d(s):
    """Quadratic Lagrange polynomials at s = -1, 0, +1 and their derivatives."""
    L = np.stack([0.5 * s * (s - 1.0), 1.0 - s * s, 0.5 * s * (s + 1.0)])
    dL = np.stack([s - 0.5, -2.0 * s, s + 0.5])
    return L, dL


# tensor-product node -> (index of the xi polynomial, index of the eta polynomial); quadratic polynomial 0 at
# -1, 1 at 0, 2 at +1
_Q9_TENSOR = [(0, 0), (2, 0), (2, 2), (0, 2), (1, 0), (2, 1), (1, 2), (0, 1), (1, 1)]
_Q4_TENSOR = [(0, 0), (1, 0), (1, 1), (0, 1)]
_CORNERS = np.array([(-1, -1), (1, -1), (1, 1), (-1, 1)], dtype=np.float64)


def _quad8(xi, eta):
    """Serendipity Q8 values and (d/dxi, d/deta): corners 1/4 (1+a xi)(1+b eta)(a xi + b eta - 1) for the
    corner (a, b); mid-sides 1/2 (1 - xi^2)(1 -+ eta) and 1/2 (1 +- xi)(1 - eta^2)."""
    N, Dx, De = [], [], []
    for a, b in _CORNERS:
        N.append(0.25 * (1 + a * xi) * (1 + b * eta) * (a * xi + b * eta - 1))
        Dx.append(0.25 * a * (1 + b * eta) * (2 * a * xi + b * eta))
        De.append(0.25 * b * (1 + a * xi) * (a * xi + 2 * b * eta))
    one = np.ones_like(xi)
    N += [0.5 * (1 - xi ** 2) * (1 - eta), 0.5 * (1 + xi) * (1 - eta ** 2),
          0.5 * (1 - xi ** 2) * (1 + eta), 0.5 * (1 - xi) * (1 - eta ** 2)]
    Dx += [-xi * (1 - eta), 0.5 * (1 - eta ** 2) * one, -xi * (1 + eta), -0.5 * (1 - eta ** 2) * one]
    De += [-0.5 * (1 - xi ** 2) * one, -(1 + xi) * eta, 0.5 * (1 - xi ** 2) * one, -(1 - xi) * eta]
    return np.stack(N, -1), np.stack([np.stack(Dx, -1), np.stack(De, -1)], -1)


def _tri3(xi, eta):
    """Barycentric coordinates (1 - xi - eta, xi, eta) and their constant gradients."""
    one, zero = np.ones_like(xi), np.zeros_like(xi)
    N = np.stack([1.0 - xi - eta, xi, eta], -1)
    D = np.stack([np.stack([-one, one, zero], -1), np.stack([-one, zero, one], -1)], -1)
    return N, D


def _tri6(xi, eta):
    """Quadratic triangle in barycentric l1 = 1 - xi - eta, l2 = xi, l3 = eta: corners l (2 l - 1),
    mid-sides 4 l1 l2, 4 l2 l3, 4 l3 l1."""
    l1, l2, l3 = 1.0 - xi - eta, xi, eta
    zero = np.zeros_like(xi)
    N = np.stack([l1 * (2 * l1 - 1), l2 * (2 * l2 - 1), l3 * (2 * l3 - 1), 4 * l1 * l2, 4 * l2 * l3, 4 * l3 * l1], -1)
    Dx = np.stack([1 - 4 * l1, 4 * l2 - 1, zero, 4 * (l1 - l2), 4 * l3, -4 * l3], -1)
    De = np.stack([1 - 4 * l1, zero, 4 * l3 - 1, -4 * l2, 4 * l2, 4 * (l1 - l3)], -1)
    return N, np.stack([Dx, De], -1)


def _tensor(element_type, xi, eta):
    if element_type == "quad4":
        (Lx, dLx), (Le, dLe), tab = _lagrange1_1d(xi), _lagrange1_1d(eta), _Q4_TENSOR
    else:
        (Lx, dLx), (Le, dLe), tab = _lagrange2_1d(xi), _lagrange2_1d(eta), _Q9_TENSOR
    N = np.stack([Lx[i] * Le[j] for i, j in tab], axis=-1)
    d_xi = np.stack([dLx[i] * Le[j] for i, j in tab], axis=-1)
    d_eta = np.stack([Lx[i] * dLe[j] for i, j in tab], axis=-1)
    return N, np.stack([d_xi, d_eta], axis=-1)


def _evaluate(element_type, xi, eta):
    _check(element_type)
    xi, eta = np.asarray(xi, np.float64), np.asarray(eta, np.float64)
    if element_type in ("quad4", "quad9"):
        return _tensor(element_type, xi, eta)
    return {"quad8": _quad8, "tri3": _tri3, "tri6": _tri6}[element_type](xi, eta)


def shape_values(element_type, xi, eta):
    """N_a(xi, eta) for arrays xi, eta of equal shape -> (..., nen)."""
    return _evaluate(element_type, xi, eta)[0]


def shape_gradients(element_type, xi, eta):
    """dN_a/d(xi, eta) for arrays xi, eta -> (..., nen, 2)."""
    return _evaluate(element_type, xi, eta)[1]


def reference_gradient_table(element_type, GP=None):
    """dN_a/dxi_j at every Gauss point: (n_q, nen, 2) float64 (the kernels' small constant table)."""
    if GP is None:
        GP, _ = element_quadrature(element_type)
    GP = np.asarray(GP, np.float64)
    return shape_gradients(element_type, GP[:, 0], GP[:, 1])


def default_nq(element_type):
    """Gauss points of the element type's default rule (the kernels are instantiated for these)."""
    return element_quadrature(element_type)[0].shape[0]


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
    _check(element_type)
    return ORDER[element_type]
