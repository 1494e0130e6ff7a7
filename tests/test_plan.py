"""Mesh plans, checked on the CPU: table consistency, colouring, and a numpy emulation of the chunked,
colour-ordered assembly the kernels perform, compared with the oracle's scatter-add product."""
import copy

import numpy as np
import pytest
import torch

from helpers import cantilever, state, unit_square
from oracle import fem
from difffea_amd import _lib
from difffea_amd.plan import CUS, NXCD, MeshPlan, jones_plassmann, tail_chunk_sizes
from difffea_amd.problem import build_problem


def _plan(p, **kw):
    return MeshPlan(p["nodes"].numpy(), p["conn"].numpy(), p["fixed"].numpy(),
                    cell_ij=None if "cell_ij" not in p else p["cell_ij"].numpy(), **kw)


def _ogrid(n_rad=6, n_circ=16):
    return build_problem(dict(element_type="quad9", geometry="plate_hole_ogrid",
                              geom_params=dict(width=1.0, radius=0.2, n_rad=n_rad, n_circ=n_circ),
                              E=1.0, nu=0.3, load_mode="force", load_value=(0.1, 0.0)))


def _unstructured(et, h):
    """Delaunay plate with a hole (irregular node valence, no cell structure)."""
    return build_problem(dict(element_type=et, mesher="unstructured", geometry="plate_hole",
                              geom_params=dict(width=1.0, height=1.0, radius=0.2), h=h, E=1.0, nu=0.3,
                              load_mode="force", load_value=(0.1, 0.0)))


MESHES = {
    "q4_37": lambda: unit_square("quad4", 37),
    "q9_19": lambda: unit_square("quad9", 19),
    "q4_cant": lambda: cantilever("quad4", h=0.05),
    "q9_ogrid": lambda: _ogrid(),
    "q9_ogrid_generic": lambda: {k: v for k, v in _ogrid().items() if k != "cell_ij"},
    "t3_30": lambda: unit_square("tri3", 30),
    "t6_13": lambda: unit_square("tri6", 13),
    "q8_17": lambda: unit_square("quad8", 17),
    "t6_cant": lambda: cantilever("tri6", h=0.1),
    "q4_unstructured": lambda: _unstructured("quad4", 0.03),
    "t3_unstructured": lambda: _unstructured("tri3", 0.03),
    "q9_unstructured": lambda: _unstructured("quad9", 0.06),
}


def _emulate_apply(plan, C, dN, v, dots=None):
    """numpy emulation of k_apply (+ k_apply_fin): per chunk gather -> core and halo elements ->
    colour-ordered LDS accumulation -> owned write / shared partial; then shared nodes sum their partials
    in slot order (partials plans).  ``dots``: a list that receives each chunk's PCG partial of v.Av (owned
    nodes v.y, shared nodes v . the chunk's partial on free components, fixed components v^2 from the chunk
    holding the node's first slot)."""
    N = plan.n_nodes
    y = np.full((N, 2), np.nan)
    part = np.zeros((plan.n_part, 2))
    vv = v.reshape(N, 2)
    shmap = plan.shmap
    inv = plan.inverse_map() if plan.mode == "partials" else None
    ls = getattr(plan, "lstride", 0)
    for c in range(plan.n_chunks):
        e0, ne, k0, nk, nown, pbase, ncol, h0, nh = plan.chunks[c, :9]
        if ls:      # fixed-stride tables: the kernels address entry k of chunk c as c * lstride + k
            assert k0 == c * ls and pbase == c * ls + nown and nk <= ls
        ln = plan.lnode[k0:k0 + nk]
        nodes = (ln & _lib.NODE_MASK).astype(np.int64)
        bits = ln >> 30
        raw = vv[nodes]
        masked = raw.copy()
        masked[bits & 1 == 1, 0] = 0
        masked[bits & 2 == 2, 1] = 0
        els = np.r_[np.arange(e0, e0 + ne), plan.hel[h0:h0 + nh]].astype(np.int64)
        loc = np.r_[plan.lelem[:, e0:e0 + ne].T, plan.hlelem[:, h0:h0 + nh].T].astype(np.int64)   # (threads, nen)
        ve = masked[loc]
        d = dN[els]
        g = np.einsum("eai,eqaj->eqij", ve, d).reshape(len(els), -1, 4)
        s_ = np.einsum("eqmn,eqn->eqm", C[els], g).reshape(len(els), -1, 2, 2)
        fe = np.einsum("eqij,eqaj->eai", s_, d)
        acc = np.zeros((nk, 2))
        col = plan.chunks[c, _lib.COL_OFF:_lib.COL_OFF + ncol + 1]
        hcol = plan.chunks[c, _lib.HCOL_OFF:_lib.HCOL_OFF + ncol + 1]
        for k in range(ncol):
            thr = np.r_[np.arange(col[k], col[k + 1]), ne + np.arange(hcol[k], hcol[k + 1])]
            idx = loc[thr].reshape(-1)
            assert len(np.unique(idx)) == len(idx), "two elements of one colour share a node"
            acc[idx] += fe[thr].reshape(-1, 2)
        if inv is not None:
            # the kernels' node sums through the inverse map: each node adds its element vectors in slot order,
            # which must be the colour order above, bit for bit
            inv_off, inv_slot = inv
            base = e0 * plan.nen
            fe_flat = fe[:ne].reshape(-1, 2)
            acc_inv = np.zeros((nk, 2))
            for k in range(nk):
                for j in range(inv_off[k0 + k], inv_off[k0 + k + 1]):
                    acc_inv[k] += fe_flat[inv_slot[j]]
                    assert inv_slot[j] // plan.nen < ne and j - base < ne * plan.nen
            assert np.array_equal(acc_inv, acc), "inverse-map sums differ from the colour-ordered sums"
        out = np.where(np.stack([bits & 1, bits & 2], 1) > 0, raw, acc)
        y[nodes[:nown]] = out[:nown]
        slots = plan.pslot[pbase:pbase + nk - nown]
        if plan.n_part:
            part[slots] = acc[nown:]
        if dots is not None:
            d = float((raw[:nown] * out[:nown]).sum())
            for j in range(nk - nown if plan.n_part else 0):   # halo plans: the rest are read-only nodes
                kk = nown + j
                fx = [bool(bits[kk] & 1), bool(bits[kk] & 2)]
                first = slots[j] == plan.sh_off[shmap[nodes[kk]]]
                for comp in range(2):
                    if not fx[comp]:
                        d += raw[kk, comp] * acc[kk, comp]
                    elif first:
                        d += raw[kk, comp] ** 2
            dots.append(d)
    for i in range(plan.n_shared):
        g_ = int(plan.sh_node[i])
        n = g_ & _lib.NODE_MASK
        acc = np.zeros(2)
        for j in range(plan.sh_off[i], plan.sh_off[i + 1]):   # slot order = chunk order
            acc += part[j]
        y[n, 0] = vv[n, 0] if g_ & _lib.FIXED_X else acc[0]
        y[n, 1] = vv[n, 1] if g_ & _lib.FIXED_Y else acc[1]
    assert not np.isnan(y).any(), "a node was written by no chunk"
    return y.reshape(-1)


def _device_view(plan):
    """The plan with the local-node tables the kernels read (``MeshPlan.device_tables``): fixed-stride node,
    slot and inverse-map tables and the chunk records pointing into them."""
    inv_off, inv_slot = plan.inverse_map()
    dt = plan.device_tables(inv_off)
    v = copy.copy(plan)
    v.chunks, v.lnode, v.pslot, v.lstride = dt["chunks"], dt["lnode"], dt["pslot"], dt["lstride"]
    v.n_lnodes = len(dt["lnode"])
    v.inverse_map = lambda: (dt["inv_off"], inv_slot)
    return v, dt


def _check_device_tables(plan):
    """Entry by entry: the fixed-stride tables hold the compact tables' entries at c * lstride + k and the
    documented padding elsewhere."""
    inv_off, _ = plan.inverse_map()
    dt = plan.device_tables(inv_off)
    ls = dt["lstride"]
    assert ls % 32 == 0 and plan.chunks[:, 3].max() <= ls <= _lib.load().dfea_chunk_nodes()
    assert len(dt["lnode"]) == len(dt["pslot"]) == len(dt["prange"]) == plan.n_chunks * ls
    assert len(dt["inv_off"]) == plan.n_chunks * ls + 1 and dt["inv_off"][-1] == plan.n_el * plan.nen
    pr = plan.prange if plan.n_part else None
    for c, rec in enumerate(plan.chunks):
        e0, ne, k0, nk, nown, pbase = (int(x) for x in rec[:6])
        b = c * ls
        assert np.array_equal(dt["chunks"][c, 2:6:3], [b, b + nown]) and np.array_equal(
            np.delete(dt["chunks"][c], [2, 5]), np.delete(rec, [2, 5]))
        assert np.array_equal(dt["lnode"][b:b + nk], plan.lnode[k0:k0 + nk])
        assert (dt["lnode"][b + nk:b + ls] == plan.lnode[k0 + nk - 1]).all()
        assert (dt["pslot"][b:b + nown] == 0).all() and (dt["pslot"][b + nk:b + ls] == 0).all()
        assert (dt["prange"][b:b + nown] == 0).all() and (dt["prange"][b + nk:b + ls] == 0).all()
        if plan.n_part:
            assert np.array_equal(dt["pslot"][b + nown:b + nk], plan.pslot[pbase:pbase + nk - nown])
            assert np.array_equal(dt["prange"][b + nown:b + nk], pr[pbase:pbase + nk - nown])
        assert np.array_equal(dt["inv_off"][b:b + nk + 1], inv_off[k0:k0 + nk + 1])
        assert (dt["inv_off"][b + nk:b + ls + 1] == (e0 + ne) * plan.nen).all()


@pytest.mark.parametrize("mode", ["halo", "partials", "small"])
@pytest.mark.parametrize("name", list(MESHES))
def test_plan_tables_and_emulated_assembly_match_oracle(name, mode):
    p = fem.precompute_geometry(MESHES[name]())
    if mode == "small":                 # small-chunk plans: quad4 in 64-element chunks
        if p["conn"].shape[1] != 4:
            pytest.skip("small-chunk plans are for quad4")
        plan = _plan(p, chunk_elems=64)
        assert plan.chunk_elems == 64 and plan.chunks[:, 1].max() <= 64
    else:
        plan = _plan(p, mode=mode)
    if mode == "halo" and plan.mode != "halo":
        pytest.skip("halo chunks would exceed the workgroup for this element type")
    plan.validate()
    # local tables reconstruct the connectivity
    nodes_of = (plan.lnode & _lib.NODE_MASK).astype(np.int64)
    ch_of_e = np.repeat(np.arange(plan.n_chunks), plan.chunks[:, 1])
    glob = nodes_of[plan.chunks[ch_of_e, 2][None, :] + plan.lelem.astype(np.int64)]
    assert np.array_equal(glob, plan.conn_int)
    assert np.array_equal(plan.conn_int.T, p["conn"].numpy()[plan.perm])
    # every node is owned by exactly one chunk (or shared, with a partial in each touching chunk)
    if plan.mode == "partials":
        assert plan.n_lnodes - plan.n_part + plan.n_shared == plan.n_nodes
    else:
        assert plan.chunks[:, 4].sum() == plan.n_nodes
    th = torch.ones(p["conn"].shape[0], dtype=torch.float64)
    u, v = state(p, 2, scale=1e-3)
    C, dN = fem.qp_setup(u, th, p)
    dots = []
    y = _emulate_apply(plan, C.numpy()[plan.perm], dN.numpy()[plan.perm], v.numpy(), dots)
    ref = fem.make_A_operator_qp(u, th, p, C=C)(v).numpy()
    assert np.abs(y - ref).max() < 1e-12 * np.abs(ref).max()
    if plan.mode == "partials":
        # the same assembly through the fixed-stride tables the kernels read gives the same bits
        _check_device_tables(plan)
        dplan, _ = _device_view(plan)
        ddots = []
        yd = _emulate_apply(dplan, C.numpy()[plan.perm], dN.numpy()[plan.perm], v.numpy(), ddots)
        assert np.array_equal(yd, y) and ddots == dots
    # the PCG's p.Ap from chunk partials alone (shared nodes by linearity) equals v.A(v)
    vAv = float(v.numpy() @ ref)
    assert abs(sum(dots) - vAv) < 1e-12 * float(np.abs(v.numpy() * ref).sum())
    # the update kernel's reconstruction of A v on shared nodes from shmap / sh_off / nfix
    sm, ss = plan.shmap, plan.shslot
    assert (sm >= 0).sum() == plan.n_shared and (ss[:, 1] > 0).sum() == plan.n_shared
    assert (ss[sm < 0] == 0).all()
    for nd in np.flatnonzero(sm >= 0):
        s_ = sm[nd]
        assert plan.sh_node[s_] & _lib.NODE_MASK == nd
        assert tuple(ss[nd]) == (plan.sh_off[s_], plan.sh_off[s_ + 1] - plan.sh_off[s_])
    # the PCG kernel's per-entry copy of the slot ranges: entry -> its node's range, which holds its slot
    if plan.n_part:
        pr = plan.prange
        nodes_of_entry = np.concatenate([(plan.lnode[c[2] + c[4]:c[2] + c[3]] & _lib.NODE_MASK).astype(np.int64)
                                         for c in plan.chunks])
        assert np.array_equal(pr, ss[nodes_of_entry][np.argsort(np.concatenate(
            [np.arange(c[5], c[5] + c[3] - c[4]) for c in plan.chunks]))])
        assert ((plan.pslot >= pr[:, 0]) & (plan.pslot < pr[:, 0] + pr[:, 1])).all()


def test_structured_meshes_get_square_tiles_and_four_colours():
    plan = _plan(unit_square("quad4", 400), balance=False)
    st = plan.stats()
    assert st["mode"] == "partials" and st["n_chunks"] == 625 and st["max_chunk_nodes"] == 289
    assert st["max_colours"] == 4 and st["max_threads"] == 256 and not plan.tail
    # tail-balanced (the default here): 512 16x16 tiles (17x17 nodes) and 226 16x8 half tiles (17x9 nodes)
    tp = _plan(unit_square("quad4", 400))
    ne, nk = tp.chunks[:, 1], tp.chunks[:, 3]
    assert tp.tail and tp.n_chunks == 738 and (ne == 256).sum() == 512 and (ne == 128).sum() == 226
    assert set(nk[ne == 256]) == {289} and set(nk[ne == 128]) == {153}
    assert tp.stats()["max_colours"] == 4 and np.array_equal(tp.chunks[:, 0], np.cumsum(ne) - ne)
    halo = _plan(unit_square("quad4", 400), mode="halo").stats()
    # 16x16 core tiles + 33 halo elements; 16x16 owned nodes + the halo's nodes (18x18)
    assert halo["mode"] == "halo" and halo["n_chunks"] == 625 and halo["max_threads"] == 289
    assert halo["max_chunk_nodes"] == 324 and halo["n_shared"] == 0
    plan9 = _plan(unit_square("quad9", 200))
    assert plan9.stats()["mode"] == "partials" and plan9.stats()["max_chunk_nodes"] == 289


def _xcd_remap(b, n):
    # the kernels' workgroup -> chunk map (workgroup b runs on XCD b mod 8; each XCD gets a contiguous range)
    x, pos = b % NXCD, b // NXCD
    base, rem = divmod(n, NXCD)
    return x * base + min(x, rem) + pos


@pytest.mark.parametrize("n_el", [160000, 144400, 90000, 80089, 160178])
def test_tail_plans_put_the_half_chunks_on_the_workgroups_past_the_whole_rounds(n_el):
    E = 256
    sizes = tail_chunk_sizes(n_el, E)
    n = len(sizes)
    rounds = n_el // (CUS * E)
    assert sizes.sum() == n_el and (sizes == E).sum() == rounds * CUS and (sizes[sizes != E] <= E // 2).all()
    assert n - rounds * CUS <= CUS                            # at most one half chunk per CU
    chunk = np.array([_xcd_remap(b, n) for b in range(n)])
    assert np.array_equal(np.sort(chunk), np.arange(n))     # the remap is a permutation
    # workgroup b shares CU b mod 256 with b +- 256: the whole rounds carry full chunks, the rest half chunks
    assert (sizes[chunk[:rounds * CUS]] == E).all() and (sizes[chunk[rounds * CUS:]] < E).all()


@pytest.mark.parametrize("et", ["quad4", "tri3"])
def test_tail_plan_emulated_assembly_matches_oracle(et, monkeypatch):
    # a GPU of 8 CUs: a 50x50 mesh (2,500 quad4 / 5,000 tri3 elements) then fills whole rounds of 256-element
    # chunks plus half chunks, so the emulation covers a tail plan at a size the oracle evaluates quickly
    import difffea_amd.plan as plan_mod
    monkeypatch.setattr(plan_mod, "CUS", 8)
    p = fem.precompute_geometry(unit_square(et, 50 if et == "quad4" else 35))
    plan = _plan(p)
    assert plan.tail and len(set(plan.chunks[:, 1])) > 1
    plan.validate(None, plan.inverse_map())
    th = torch.ones(p["conn"].shape[0], dtype=torch.float64)
    u, v = state(p, 4, scale=1e-3)
    C, dN = fem.qp_setup(u, th, p)
    dots = []
    y = _emulate_apply(plan, C.numpy()[plan.perm], dN.numpy()[plan.perm], v.numpy(), dots)
    ref = fem.make_A_operator_qp(u, th, p, C=C)(v).numpy()
    assert np.abs(y - ref).max() < 1e-12 * np.abs(ref).max()
    if plan.mode == "partials":
        # the same assembly through the fixed-stride tables the kernels read gives the same bits
        _check_device_tables(plan)
        dplan, _ = _device_view(plan)
        ddots = []
        yd = _emulate_apply(dplan, C.numpy()[plan.perm], dN.numpy()[plan.perm], v.numpy(), ddots)
        assert np.array_equal(yd, y) and ddots == dots
    vAv = float(v.numpy() @ ref)
    assert abs(sum(dots) - vAv) < 1e-12 * float(np.abs(v.numpy() * ref).sum())


def test_tail_plans_only_where_they_balance():
    assert tail_chunk_sizes(65536, 256) is None            # exactly one round
    assert tail_chunk_sizes(50000, 256) is None            # under one round: one chunk per CU or fewer
    assert tail_chunk_sizes(122500, 256) is None           # 478 chunks: remainder past half a round
    assert tail_chunk_sizes(566 * 566, 256) is None        # more than three rounds
    p = unit_square("quad4", 300)                          # 90,000 elements: 1 round + 96 half chunks
    a, b = _plan(p), _plan(p, balance=False)
    assert a.tail and not b.tail and a.n_chunks == 448 and b.n_chunks == 352
    # the same elements, a different cut
    assert np.array_equal(np.sort(a.perm), np.sort(b.perm))


def test_jones_plassmann_is_a_valid_colouring():
    conn = _ogrid(4, 12)["conn"].numpy()
    col = jones_plassmann(conn)
    for n in range(conn.max() + 1):
        els = np.flatnonzero((conn == n).any(1))
        assert len(np.unique(col[els])) == len(els)
    assert col.max() < _lib.MAXCOL


def test_plan_rejects_oversized_chunks():
    p = unit_square("quad4", 40)
    with pytest.raises(ValueError):
        _plan(p, max_nodes=100)


def _corrupt_cases():
    """(name, mutate(plan, inv) -> None): each breaks one index the kernels follow."""
    def set_(attr, idx, val):
        def f(pl, inv):
            a = getattr(pl, attr)
            a[idx] = val
        return f

    def inv_off_(idx, val):
        def f(pl, inv):
            inv[0][idx] = val
        return f

    def inv_slot_(idx, val):
        def f(pl, inv):
            inv[1][idx] = val
        return f
    return [
        ("chunk element range", set_("chunks", (0, 1), 10 ** 6)),
        ("chunk node count over the LDS capacity", set_("chunks", (0, 3), 700)),
        ("element local node index", set_("lelem", (0, 0), 999)),
        ("local node id", set_("lnode", 0, 0x3FFFFFF0)),
        ("partial slot", set_("pslot", 0, 10 ** 6)),
        ("shared slot ranges", set_("sh_off", -1, 10 ** 6)),
        ("connectivity", set_("conn_int", (0, 0), -1)),
        ("inverse map offsets", inv_off_(3, 10 ** 6)),
        ("inverse map slot", inv_slot_(5, 4000)),
    ]


@pytest.mark.parametrize("case", range(9))
def test_validate_rejects_corrupt_tables(case):
    """MeshPlan.validate raises ValueError (not assert: it must survive python -O) for every corrupted
    table the kernels would follow without checks."""
    import copy
    p = unit_square("quad4", 40)
    base = MeshPlan(p["nodes"].numpy(), p["conn"].numpy(), p["fixed"].numpy())
    inv0 = base.inverse_map()
    base.validate(640, inv0)
    name, mutate = _corrupt_cases()[case]
    pl = copy.deepcopy(base)
    inv = (inv0[0].copy(), inv0[1].copy())
    mutate(pl, inv)
    with pytest.raises(ValueError):
        pl.validate(640, inv)


def test_chunking_choice():
    """quad4 meshes up to SMALL_QUAD4_MAX_ELEMS elements get 64-element chunks (the DFEA_SMALL kernels),
    larger ones and other element types the element type's default."""
    from difffea_amd.plan import SMALL_QUAD4_MAX_ELEMS, chunking_for
    assert chunking_for(4, 10000) == (64, _lib.SMALL) and chunking_for(4, SMALL_QUAD4_MAX_ELEMS)[1] == _lib.SMALL
    assert chunking_for(4, SMALL_QUAD4_MAX_ELEMS + 1) == (None, 0)
    assert chunking_for(4, 10 ** 6, "small") == (64, _lib.SMALL)
    assert chunking_for(4, 100, "large") == (None, 0)
    assert chunking_for(9, 100) == (None, 0)
    with pytest.raises(ValueError):
        chunking_for(4, 100, "tiny")
