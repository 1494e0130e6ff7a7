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

All functions return ``(nodes (N,2) float64 tensor, conn (n_el,nen) int64 tensor)``.
"""
import math

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
