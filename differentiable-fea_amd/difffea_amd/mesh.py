"""Structured meshes for the hot-path workloads (host-side set-up, CPU).

* ``create_2d_quad_mesh`` / ``create_2d_quad9_mesh`` / ``create_2d_quad8_mesh`` / ``create_2d_tri3_mesh`` /
  ``create_2d_tri6_mesh``: rectangle [0,Lx]x[0,Ly] with the reference's node numbering (x-major node ids
  ``i*(ny+1)+j``, elements ``i*ny+j``, CCW corners; Q9 on the stride-2 grid, corners -> mid-sides -> centre;
  Q8 the same grid without the cell centres, renumbered consecutively; triangles split every cell along the
  diagonal from its top-left to its bottom-right corner, all lower-left triangles first, then all upper-right
  ones), `src/difffea/mesh.py:19-421`.  Coordinates come from ``torch.linspace`` exactly as in the reference
  so that node coordinates agree bit for bit.
* ``create_plate_hole_ogrid_q9``: square [-w/2,w/2]^2 with a centred circular hole as a structured O-grid
  of Q9 elements, linear blending between the circle and the square boundary along each ray.  This mesh
  has no counterpart in the reference (which meshes curved geometries with Gmsh).
* Unstructured meshes (`src/difffea/gmsh_mesh.py:41-97,135-188`): ``load_mesh_npz`` reads the mesher's
  exchange format (an ``.npz`` file with ``xy`` (N, 2) float64 and ``conn`` (n_el, nen) int64, corners first,
  counter-clockwise, then mid-sides, then the Q9 centre), drops unused nodes (``filter_unused_nodes``) and
  makes every element counter-clockwise (``check_and_fix_ccw``); ``save_mesh_npz`` writes the format.
  Gmsh itself is not part of this build; ``create_unstructured_plate_hole`` is a Gmsh-free mesher (Delaunay
  triangles of jittered points, optionally split into quadrilaterals) that produces unstructured meshes of
  every element type with irregular node valence.

All functions return ``(nodes (N,2) float64 tensor, conn (n_el,nen) int64 tensor)``.
"""
import math

import numpy as np
import torch


def create_2d_quad_mesh(Lx, Ly, nx, ny, dtype=torch.float64):
    """Q4 mesh of [0,Lx]x[0,Ly] with nx x ny cells."""
    x = torch.linspace(0.0, Lx, nx + 1, dtype=dtype)
    y = torch.linspace(0.0, Ly, ny + 1, dtype=dtype)
    gx, gy = torch.meshgrid(x, y, indexing="ij")
    nodes = torch.stack([gx.flatten(), gy.flatten()], dim=-1)
    grid = torch.arange((nx + 1) * (ny + 1)).reshape(nx + 1, ny + 1)
    conn = torch.stack([grid[:-1, :-1].flatten(), grid[1:, :-1].flatten(),
                        grid[1:, 1:].flatten(), grid[:-1, 1:].flatten()], dim=-1)
    return nodes, conn


def create_2d_quad9_mesh(Lx, Ly, nx, ny, dtype=torch.float64):
    """Q9 mesh of [0,Lx]x[0,Ly] with nx x ny cells (nodes on the (2nx+1)x(2ny+1) grid)."""
    gnx, gny = 2 * nx, 2 * ny
    nodes, _ = create_2d_quad_mesh(Lx, Ly, gnx, gny, dtype=dtype)
    g = torch.arange((gnx + 1) * (gny + 1)).reshape(gnx + 1, gny + 1)

    def s(i0, j0):                         # node at grid offset (i0, j0) of every cell, cells x-major
        return g[i0:i0 + gnx - 1:2, j0:j0 + gny - 1:2].flatten()

    conn = torch.stack([s(0, 0), s(2, 0), s(2, 2), s(0, 2),      # corners (CCW)
                        s(1, 0), s(2, 1), s(1, 2), s(0, 1),      # mid-sides: bottom, right, top, left
                        s(1, 1)], dim=-1)                        # centre
    return nodes, conn


def create_2d_quad8_mesh(Lx, Ly, nx, ny, dtype=torch.float64):
    """Serendipity Q8 mesh: the Q9 node grid without the cell centres (odd, odd grid points), nodes renumbered
    consecutively in grid order; element nodes = corners (CCW), then mid-sides bottom, right, top, left."""
    nodes9, conn9 = create_2d_quad9_mesh(Lx, Ly, nx, ny, dtype=dtype)
    gnx, gny = 2 * nx, 2 * ny
    keep = torch.ones(gnx + 1, gny + 1, dtype=torch.bool)
    keep[1:gnx:2, 1:gny:2] = False
    keep = keep.flatten()
    new_id = torch.full((keep.numel(),), -1, dtype=torch.int64)
    new_id[keep] = torch.arange(int(keep.sum()))
    return nodes9[keep], new_id[conn9[:, :8]]


def create_2d_tri3_mesh(Lx, Ly, nx, ny, dtype=torch.float64):
    """Tri3 mesh: each Q4 cell (n0, n1, n2, n3) CCW from bottom-left is cut along n3-n1 into the lower-left
    triangle (n0, n1, n3) and the upper-right triangle (n1, n2, n3), both CCW; all lower-left triangles (in cell
    order) come first, then all upper-right ones."""
    nodes, q = create_2d_quad_mesh(Lx, Ly, nx, ny, dtype=dtype)
    lower = q[:, [0, 1, 3]]
    upper = q[:, [1, 2, 3]]
    return nodes, torch.cat([lower, upper], dim=0)


def create_2d_tri6_mesh(Lx, Ly, nx, ny, dtype=torch.float64):
    """Tri6 mesh: each Q9 cell cut along its top-left / bottom-right diagonal, whose mid-point is the cell
    centre.  Lower-left triangle: corners (1, 2, 4), mid-sides (5, 9, 8); upper-right: corners (2, 3, 4),
    mid-sides (6, 7, 9) in the Q9 cell's 1-based node numbers; lower-left triangles first."""
    nodes, q = create_2d_quad9_mesh(Lx, Ly, nx, ny, dtype=dtype)
    lower = q[:, [0, 1, 3, 4, 8, 7]]
    upper = q[:, [1, 2, 3, 5, 6, 8]]
    return nodes, torch.cat([lower, upper], dim=0)


_BUILDERS = {"quad4": create_2d_quad_mesh, "quad8": create_2d_quad8_mesh, "quad9": create_2d_quad9_mesh,
             "tri3": create_2d_tri3_mesh, "tri6": create_2d_tri6_mesh}


def create_rectangle_mesh(element_type, Lx, Ly, nx, ny, dtype=torch.float64):
    """Structured mesh of [0,Lx]x[0,Ly] with nx x ny cells of any element type; triangle meshes have 2 nx ny
    elements (`src/difffea/mesh.py:424-438`)."""
    if element_type not in _BUILDERS:
        raise ValueError(f"Unknown element type '{element_type}'. Choose one of {list(_BUILDERS)}.")
    return _BUILDERS[element_type](Lx, Ly, nx, ny, dtype=dtype)


def create_plate_hole_ogrid_q9(width=1.0, radius=0.2, n_rad=64, n_circ=256, dtype=torch.float64):
    """O-grid of Q9 elements around a centred hole.

    Parametric node grid (s, t): s = k/(2 n_rad) radial from the hole (s=0) to the outer square (s=1),
    t = l/(2 n_circ) counter-clockwise angle fraction (periodic).  A point is
    X = (1-s) r (cos phi, sin phi) + s (w/2)(cos phi, sin phi)/max(|cos phi|, |sin phi|), phi = 2 pi t,
    so hole-edge nodes lie exactly on the circle and outer nodes exactly on the square.  Mid-side and
    centre nodes sit at parametric midpoints.  Local xi runs radially outward and eta counter-clockwise,
    which makes every element right-handed (det J > 0).

    Node id = l * (2 n_rad + 1) + k.  Element id = ic * n_rad + ir.
    Returns (nodes, conn, cell_ij) with cell_ij (n_el, 2) = (ir, ic), the logical cell indices.
    """
    nk, nl = 2 * n_rad + 1, 2 * n_circ
    k = torch.arange(nk, dtype=dtype)
    l = torch.arange(nl, dtype=dtype)
    s = k / (2 * n_rad)
    phi = 2.0 * math.pi * l / (2 * n_circ)
    c, sn = torch.cos(phi), torch.sin(phi)
    # snap the trigonometric values of the axis and diagonal directions so the square edges are exact
    inner = torch.stack([radius * c, radius * sn], dim=-1)                       # (nl, 2)
    outer = 0.5 * width * torch.stack([c, sn], dim=-1) / torch.maximum(c.abs(), sn.abs())[:, None]
    outer = torch.where(outer.abs() > 0.5 * width * (1 - 1e-13), torch.sign(outer) * 0.5 * width, outer)
    X = (1.0 - s)[None, :, None] * inner[:, None, :] + s[None, :, None] * outer[:, None, :]   # (nl, nk, 2)
    nodes = X.reshape(-1, 2).contiguous()

    def nid(kk, ll):
        return (ll % nl) * nk + kk

    ir = torch.arange(n_rad)
    ic = torch.arange(n_circ)
    IC, IR = torch.meshgrid(ic, ir, indexing="ij")
    k0, l0 = (2 * IR).flatten(), (2 * IC).flatten()
    conn = torch.stack([nid(k0, l0), nid(k0 + 2, l0), nid(k0 + 2, l0 + 2), nid(k0, l0 + 2),
                        nid(k0 + 1, l0), nid(k0 + 2, l0 + 1), nid(k0 + 1, l0 + 2), nid(k0, l0 + 1),
                        nid(k0 + 1, l0 + 1)], dim=-1)
    cell_ij = torch.stack([IR.flatten(), IC.flatten()], dim=-1)
    return nodes, conn, cell_ij


# ------------------------------------------------------------------------------------------------------------
# unstructured meshes: exchange format, clean-up, a Gmsh-free mesher
# ------------------------------------------------------------------------------------------------------------
N_CORNERS = {"quad4": 4, "quad8": 4, "quad9": 4, "tri3": 3, "tri6": 3}
NEN = {"quad4": 4, "quad8": 8, "quad9": 9, "tri3": 3, "tri6": 6}


def filter_unused_nodes(nodes, conn):
    """Drop nodes no element uses and renumber the connectivity consecutively, keeping the nodes' order
    (`src/difffea/gmsh_mesh.py:41-58`).  Returns (nodes (N', 2), conn (n_el, nen))."""
    used, inverse = torch.unique(conn, return_inverse=True)
    return nodes[used], inverse.reshape(conn.shape)


def signed_corner_area(nodes, conn, n_corners):
    """Shoelace area of the corner polygon of every element, > 0 counter-clockwise (`gmsh_mesh.py:61-75`)."""
    xy = nodes[conn[:, :n_corners]]
    x, y = xy[..., 0], xy[..., 1]
    return 0.5 * torch.sum(x * torch.roll(y, -1, 1) - torch.roll(x, -1, 1) * y, dim=1)


def _reversal(element_type):
    """Local-node permutation that reverses an element's orientation and keeps node 0: the corners run
    backwards, each mid-side node follows its edge (edge k of the reversed element is edge n_c - 1 - k of the
    original), the centre stays (`gmsh_mesh.py:79-86`)."""
    nc, nen = N_CORNERS[element_type], NEN[element_type]
    corners = [0] + list(range(nc - 1, 0, -1))
    mids = [nc + (nc - 1 - k) for k in range(nc)] if nen > nc else []
    return corners + mids + list(range(2 * nc, nen))


def check_and_fix_ccw(nodes, conn, element_type):
    """Reorder the local nodes of every clockwise element (negative corner area) so that all elements are
    counter-clockwise, i.e. det J > 0 (`src/difffea/gmsh_mesh.py:89-103`)."""
    area = signed_corner_area(nodes, conn, N_CORNERS[element_type])
    flipped = conn[:, _reversal(element_type)]
    return torch.where((area < 0)[:, None], flipped, conn)


def load_mesh_npz(path, element_type, dtype=torch.float64):
    """Read a mesh in the mesher's ``.npz`` exchange format (``xy``, ``conn``), drop unused nodes and make
    every element counter-clockwise.  Nothing is unpickled (``allow_pickle=False``)."""
    if element_type not in NEN:
        raise ValueError(f"Unknown element type '{element_type}'. Choose one of {list(NEN)}.")
    with np.load(path, allow_pickle=False) as z:
        if "xy" not in z or "conn" not in z:
            raise ValueError(f"{path}: a mesh file holds the arrays 'xy' and 'conn'")
        xy, conn = np.asarray(z["xy"]), np.asarray(z["conn"])
    if xy.ndim != 2 or xy.shape[1] < 2 or conn.ndim != 2 or conn.shape[1] != NEN[element_type]:
        raise ValueError(f"{path}: xy must be (N, 2) and conn (n_el, {NEN[element_type]}) for {element_type}, "
                         f"got {xy.shape} and {conn.shape}")
    if not np.issubdtype(conn.dtype, np.integer) or conn.size == 0:
        raise ValueError(f"{path}: conn must be a non-empty integer array")
    if conn.min() < 0 or conn.max() >= xy.shape[0]:
        raise ValueError(f"{path}: conn refers to nodes outside 0..{xy.shape[0] - 1}")
    nodes = torch.tensor(xy[:, :2], dtype=dtype)
    nodes, conn_t = filter_unused_nodes(nodes, torch.tensor(conn, dtype=torch.int64))
    return nodes, check_and_fix_ccw(nodes, conn_t, element_type)


def save_mesh_npz(path, nodes, conn):
    """Write ``nodes`` / ``conn`` in the ``.npz`` exchange format (``xy`` float64, ``conn`` int64)."""
    np.savez(path, xy=np.asarray(nodes, dtype=np.float64), conn=np.asarray(conn, dtype=np.int64))


def _jittered_points(width, height, radius, h, rng):
    """Boundary points at spacing ~h on the rectangle [0,W]x[0,H] and on the hole (centre of the plate),
    interior points on a jittered grid at least h/2 from both."""
    W, H = width, height
    cx, cy = 0.5 * W, 0.5 * H
    pts = []
    for a, b, n in (((0, 0), (W, 0), max(2, round(W / h))), ((W, 0), (W, H), max(2, round(H / h))),
                    ((W, H), (0, H), max(2, round(W / h))), ((0, H), (0, 0), max(2, round(H / h)))):
        t = np.arange(n) / n
        pts.append(np.stack([a[0] + t * (b[0] - a[0]), a[1] + t * (b[1] - a[1])], 1))
    if radius > 0:
        n = max(8, round(2 * math.pi * radius / h))
        phi = 2 * math.pi * np.arange(n) / n
        pts.append(np.stack([cx + radius * np.cos(phi), cy + radius * np.sin(phi)], 1))
    gx, gy = np.meshgrid(np.arange(h, W - 0.5 * h, h), np.arange(h, H - 0.5 * h, h), indexing="ij")
    q = np.stack([gx.ravel(), gy.ravel()], 1) + rng.uniform(-0.3 * h, 0.3 * h, (gx.size, 2))
    inside = (q[:, 0] > 0.5 * h) & (q[:, 0] < W - 0.5 * h) & (q[:, 1] > 0.5 * h) & (q[:, 1] < H - 0.5 * h)
    if radius > 0:
        inside &= np.hypot(q[:, 0] - cx, q[:, 1] - cy) > radius + 0.5 * h
    pts.append(q[inside])
    return np.concatenate(pts)


def _edge_nodes(xy, conn, edges):
    """One new node at the midpoint of every distinct edge; returns (xy', (n_el, n_edges) new node ids)."""
    e = np.stack([conn[:, list(ed)] for ed in edges], 1)                  # (n_el, n_edges, 2)
    key = np.sort(e, axis=2).reshape(-1, 2)
    uniq, inv = np.unique(key, axis=0, return_inverse=True)
    mid = 0.5 * (xy[uniq[:, 0]] + xy[uniq[:, 1]])
    return np.concatenate([xy, mid]), (xy.shape[0] + inv.reshape(conn.shape[0], len(edges)))


def create_unstructured_plate_hole(element_type="tri3", width=1.0, height=1.0, radius=0.2, h=0.05, seed=0,
                                   dtype=torch.float64):
    """Unstructured mesh of the plate [0,W]x[0,H] with a centred hole of ``radius`` (0: no hole), target
    edge length ~h, for any element type.

    Triangles: Delaunay triangulation (scipy) of jittered points; triangles whose centroid lies in the hole are
    dropped.  Quadrilaterals: every triangle (a, b, c) is split into the three quadrilaterals
    (a, m_ab, g, m_ca), (b, m_bc, g, m_ab), (c, m_ca, g, m_bc) at its edge midpoints m and centroid g, so
    node valences vary from 3 to about 10.  Quadratic elements get mid-side nodes at the straight-edge
    midpoints (and Q9 a centre node at the mean of the corners).  Every element is counter-clockwise; the
    nodes and elements come in the order the construction makes them (no spatial sorting).
    """
    from scipy.spatial import Delaunay
    if element_type not in NEN:
        raise ValueError(f"Unknown element type '{element_type}'. Choose one of {list(NEN)}.")
    rng = np.random.default_rng(seed)
    h_tri = h * (2.0 if N_CORNERS[element_type] == 4 else 1.0)    # quads halve the triangle edges
    xy = _jittered_points(width, height, radius, h_tri, rng)
    tri = Delaunay(xy).simplices.astype(np.int64)
    g = xy[tri].mean(axis=1)
    if radius > 0:
        tri = tri[np.hypot(g[:, 0] - 0.5 * width, g[:, 1] - 0.5 * height) > radius]
    a = xy[tri]
    area = 0.5 * ((a[:, 1, 0] - a[:, 0, 0]) * (a[:, 2, 1] - a[:, 0, 1]) -
                  (a[:, 2, 0] - a[:, 0, 0]) * (a[:, 1, 1] - a[:, 0, 1]))
    tri = np.where((area < 0)[:, None], tri[:, [0, 2, 1]], tri)
    tri = tri[np.abs(area) > 1e-12 * h_tri * h_tri]
    if N_CORNERS[element_type] == 3:
        conn = tri
        if element_type == "tri6":
            xy, m = _edge_nodes(xy, conn, [(0, 1), (1, 2), (2, 0)])
            conn = np.concatenate([conn, m], 1)
    else:
        xy, m = _edge_nodes(xy, tri, [(0, 1), (1, 2), (2, 0)])          # m_ab, m_bc, m_ca
        gid = xy.shape[0] + np.arange(tri.shape[0])
        xy = np.concatenate([xy, xy[tri].mean(axis=1)])
        A, B, C = tri[:, 0], tri[:, 1], tri[:, 2]
        mab, mbc, mca = m[:, 0], m[:, 1], m[:, 2]
        conn = np.concatenate([np.stack([A, mab, gid, mca], 1), np.stack([B, mbc, gid, mab], 1),
                               np.stack([C, mca, gid, mbc], 1)])
        if element_type in ("quad8", "quad9"):
            xy, mq = _edge_nodes(xy, conn, [(0, 1), (1, 2), (2, 3), (3, 0)])
            conn = np.concatenate([conn, mq], 1)
            if element_type == "quad9":
                cid = xy.shape[0] + np.arange(conn.shape[0])
                xy = np.concatenate([xy, xy[conn[:, :4]].mean(axis=1)])
                conn = np.concatenate([conn, cid[:, None]], 1)
    nodes, conn_t = filter_unused_nodes(torch.tensor(xy, dtype=dtype), torch.tensor(conn, dtype=torch.int64))
    return nodes, check_and_fix_ccw(nodes, conn_t, element_type)
