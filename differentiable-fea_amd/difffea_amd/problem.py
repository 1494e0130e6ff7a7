"""Problem set-up: specification dict -> ``params`` dict (host side, CPU float64).

The ``params`` keys are the data contract of the reference's solver functions
(`src/difffea/operators.py:19-34`, `src/difffea/problem.py:180-187`):

    nodes (N,2) f64, conn (n_el,nen) i64, GP (n_q,2), GW (n_q,), material, mat_p (n_p,), mat_scalable (n_p,),
    f (2N,), m (2N,) {0,1}, fixed (2N,) bool, u_D (2N,), element_type

Materials: every entry of the library (``spec["material"]``, parameters by name in
``spec["material_params"]``, neo-Hookean from ``E``, ``nu`` by default).  Elements: tri3, quad4, tri6,
quad8, quad9.
Geometries:

* ``"rectangle"``: ``geom_params={"width","height"}`` and either ``nx, ny`` or ``h`` (cells ~ size/h),
  as the reference's structured mesher (`src/difffea/problem.py:95-109`);
* ``"plate_hole_ogrid"``: ``geom_params={"width","radius","n_rad","n_circ"}``, quad9 only.

Meshers (``spec["mesher"]``): ``"structured"`` (default; the geometries above), ``"unstructured"`` (Delaunay
mesher of ``mesh.create_unstructured_plate_hole`` for ``"rectangle"`` and ``"plate_hole"``
(``geom_params={"width","height","radius"}``), edge length ``h``, seed ``mesh_seed``) and ``"npz"`` (a mesh
file in the Gmsh exchange format, ``spec["mesh_file"]``, cleaned and made counter-clockwise on load).

Supports and loads follow `src/difffea/problem.py:115-194`: clamped side, prescribed displacement
(``load_mode="displacement"``, components may be ``None``) or a consistent total edge force
(``load_mode="force"``).
"""
import torch

from . import mesh as _mesh
from .boundary_conditions import (boundary_edges, dofs_from_node_mask, edge_traction_load,
                                  element_edges, select_boundary_nodes, traction_order)
from .elements import element_quadrature
from .materials import material_vector


def lame_parameters(E, nu):
    """(mu, lambda) from Young's modulus and Poisson's ratio (`src/difffea/problem.py:46-55`)."""
    mu = E / (2.0 * (1.0 + nu))
    lmbda = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    return mu, lmbda


def build_mesh(spec):
    et = spec["element_type"]
    geom = spec.get("geometry", "rectangle")
    gp = spec.get("geom_params", {})
    mesher = spec.get("mesher", "structured")
    if mesher == "npz":
        nodes, conn = _mesh.load_mesh_npz(spec["mesh_file"], et)
        return nodes, conn, None
    if mesher == "unstructured":
        if geom not in ("rectangle", "plate_hole"):
            raise ValueError("the unstructured mesher knows the geometries 'rectangle' and 'plate_hole'")
        W, H = float(gp.get("width", 1.0)), float(gp.get("height", 1.0))
        r = float(gp.get("radius", 0.2)) if geom == "plate_hole" else 0.0
        nodes, conn = _mesh.create_unstructured_plate_hole(et, W, H, r, float(spec.get("h", 0.05)),
                                                           int(spec.get("mesh_seed", 0)))
        return nodes, conn, None
    if mesher == "gmsh":
        raise ValueError("Gmsh is not part of this build: mesh with Gmsh elsewhere and load its .npz output "
                         "with mesher='npz', or use mesher='structured' / 'unstructured'")
    if mesher != "structured":
        raise ValueError(f"Unknown mesher '{mesher}'. Choose 'structured', 'unstructured' or 'npz'.")
    if geom == "rectangle":
        W, H = float(gp.get("width", 1.0)), float(gp.get("height", 1.0))
        if "nx" in spec:
            nx, ny = int(spec["nx"]), int(spec.get("ny", spec["nx"]))
        else:
            nx, ny = max(1, round(W / spec["h"])), max(1, round(H / spec["h"]))
        nodes, conn = _mesh.create_rectangle_mesh(et, W, H, nx, ny)
        return nodes, conn, None
    if geom == "plate_hole_ogrid":
        if et != "quad9":
            raise ValueError("the O-grid plate with a hole is built from quad9 elements only")
        nodes, conn, cell_ij = _mesh.create_plate_hole_ogrid_q9(
            float(gp.get("width", 1.0)), float(gp.get("radius", 0.2)), int(gp.get("n_rad", 64)),
            int(gp.get("n_circ", 256)))
        return nodes, conn, cell_ij
    raise ValueError(f"Unknown geometry '{geom}'. Choose 'rectangle' or 'plate_hole_ogrid'.")


def build_problem(spec):
    """Assemble the ``params`` dictionary from a specification (see module docstring)."""
    et = spec["element_type"]
    nodes, conn, cell_ij = build_mesh(spec)
    N = nodes.shape[0]
    GP, GW = element_quadrature(et)

    fixed_nodes = select_boundary_nodes(nodes, spec.get("fixed_side", "left"))
    comps = {"xy": (0, 1), "x": (0,), "y": (1,)}[spec.get("fixed_dofs", "xy")]
    fixed = dofs_from_node_mask(fixed_nodes, comps)
    u_D = torch.zeros(2 * N, dtype=torch.float64)
    f = torch.zeros(N, 2, dtype=torch.float64)

    load_nodes = select_boundary_nodes(nodes, spec.get("load_side", "right"))
    value = spec.get("load_value", (0.0, 0.0))
    if spec.get("load_mode", "force") == "displacement":
        for c, val in enumerate(value):
            if val is None:
                continue
            dm = dofs_from_node_mask(load_nodes, (c,))
            fixed = fixed | dm
            u_D = torch.where(dm, torch.as_tensor(float(val), dtype=torch.float64), u_D)
    else:
        edges = boundary_edges(conn, element_edges(et))
        f = edge_traction_load(nodes, edges, traction_order(et), load_nodes, (float(value[0]), float(value[1])))

    m = torch.where(fixed, 0.0, 1.0).to(torch.float64)
    # material (`src/difffea/problem.py:172-178`): neo-Hookean from E, nu unless parameters are given
    material = spec.get("material", "neo_hookean")
    values = spec.get("material_params")
    if values is None and material == "neo_hookean":
        mu0, l0 = lame_parameters(float(spec.get("E", 1.0)), float(spec.get("nu", 0.3)))
        values = {"mu": mu0, "lmbda": l0}
    mat_p, mat_scalable = material_vector(material, values)
    return {
        "nodes": nodes, "conn": conn,
        "GP": torch.from_numpy(GP), "GW": torch.from_numpy(GW),
        "material": material,
        "mat_p": mat_p,
        "mat_scalable": mat_scalable,
        "f": f.reshape(-1).contiguous(), "m": m, "fixed": fixed, "u_D": u_D,
        "element_type": et,
        **({"cell_ij": cell_ij} if cell_ij is not None else {}),
    }


def to_device(params, device):
    """Move every tensor of ``params`` to ``device`` (functional; `src/difffea/problem.py:86-92`)."""
    return {k: (v.to(device) if torch.is_tensor(v) else v) for k, v in params.items()}
