"""Mesh plan: the integer tables that let one workgroup assemble one chunk of elements deterministically.

Built once per (mesh, Dirichlet mask) on the host with vectorised numpy; uploaded to HBM once.

1. Element order.  Elements are sorted along a Morton (Z) curve of their cell coordinates, then cut into
   chunks of ``chunk_elems`` consecutive elements.  On structured grids the cell coordinates are the exact
   integer cell indices (detected by rank), so chunks are square tiles (16x16 quad4, 8x8 quad9); otherwise
   element centroids are quantised onto a power-of-two grid.  Tail-balanced plans (linear elements, see
   ``tail_chunk_sizes``) cut the elements past whole rounds of 256 chunks into chunks of half the size.
2. Colours.  Elements are coloured so that two elements of one colour never share a node: the parity of
   the cell indices on structured grids (4 colours), else Jones-Plassmann on the node-sharing graph.  Inside
   a chunk the elements are sorted by colour, so each colour is a contiguous range of threads.
3. Local nodes.  Each chunk lists its distinct nodes, owned nodes first (used by no other chunk), then
   shared nodes.  ``lelem[a][e]`` is the local index of element node a.  Fixed dofs are folded into the
   node ids as bits 30/31.
4. Shared nodes.  Every (chunk, shared node) pair gets a partial slot.  Slots are ordered by shared node,
   then chunk: the partials of shared node i are slots ``sh_off[i]..sh_off[i+1]-1``, summed in that order.
   ``pslot`` maps a chunk's shared entries (chunk order) to their slots.

Halo mode (quad4, on request).  Every node is owned by exactly one chunk: the chunk of its adjacent element
latest in Morton order (on a grid: the element up and to the right, so a 16x16 tile owns a 16x16 block of
nodes).  A chunk evaluates its core elements plus the "halo" elements of other chunks that touch its owned
nodes (33 per interior tile), so each owned node is completed inside one workgroup: no shared nodes, no
partials, no second kernel, at the price of ~13 % more element reads (on MI355X the partials plan with its
finalize kernel measured faster, so it is the default).  Halo elements are listed per chunk
(``hel``, with their local indices in ``hlelem``), sorted by colour like the core elements.  Partials mode
(the default, and the fallback when a halo chunk would exceed the workgroup) is described in 4.

Nothing here depends on which other solves run in the same batch, so results are batch-independent.
"""
import numpy as np

from . import _lib


def _part1by1(x):
    x = x.astype(np.uint64) & np.uint64(0xFFFFFFFF)
    x = (x | (x << np.uint64(16))) & np.uint64(0x0000FFFF0000FFFF)
    x = (x | (x << np.uint64(8))) & np.uint64(0x00FF00FF00FF00FF)
    x = (x | (x << np.uint64(4))) & np.uint64(0x0F0F0F0F0F0F0F0F)
    x = (x | (x << np.uint64(2))) & np.uint64(0x3333333333333333)
    x = (x | (x << np.uint64(1))) & np.uint64(0x5555555555555555)
    return x


def morton(ix, iy):
    return _part1by1(ix) | (_part1by1(iy) << np.uint64(1))


def _rank(vals, tol):
    """Rank of every value among the distinct values (values closer than tol are equal)."""
    q = np.round(vals / tol).astype(np.int64)
    u, inv = np.unique(q, return_inverse=True)
    return inv, len(u)


def cell_coordinates(nodes, conn):
    """Integer cell coordinates of the elements and whether they form an exact structured grid."""
    c = nodes[conn].mean(axis=1)
    n_el = conn.shape[0]
    ext = max(float(np.ptp(c[:, 0])), float(np.ptp(c[:, 1])), 1e-300)
    ix, nx = _rank(c[:, 0], 1e-9 * ext)
    iy, ny = _rank(c[:, 1], 1e-9 * ext)
    if nx * ny == n_el and len(np.unique(ix * ny + iy)) == n_el:
        return ix, iy, True
    bits = max(1, int(np.ceil(np.log2(max(2.0, 2.0 * np.sqrt(n_el))))))
    g = (1 << bits) - 1
    lo = c.min(axis=0)
    sc = np.ptp(c, axis=0)
    sc[sc == 0] = 1.0
    q = np.floor((c - lo) / sc * g).astype(np.int64)
    return q[:, 0], q[:, 1], False


def _parity_colouring_valid(conn, ix, iy):
    """True when every two node-sharing elements differ in the parity of ix or of iy, so that the
    4-colour parity colouring is a valid colouring (fails e.g. for periodic grids of odd period)."""
    a, b = _node_sharing_pairs(conn)
    return bool(np.all((ix[a] % 2 != ix[b] % 2) | (iy[a] % 2 != iy[b] % 2)))


def _node_sharing_pairs(conn):
    """All (e1, e2), e1 != e2, of elements that share at least one node."""
    n_el, nen = conn.shape
    e = np.repeat(np.arange(n_el), nen)
    n = conn.reshape(-1)
    order = np.argsort(n, kind="stable")
    n_s, e_s = n[order], e[order]
    starts = np.flatnonzero(np.r_[True, n_s[1:] != n_s[:-1]])
    counts = np.diff(np.r_[starts, len(n_s)])
    a_list, b_list = [], []
    for k in range(1, counts.max()):
        sel = starts[counts > k]
        # pair element at position start+j with element at start+k, for all j < k
        for j in range(k):
            a_list.append(e_s[sel + j])
            b_list.append(e_s[sel + k])
    if not a_list:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    a = np.concatenate(a_list)
    b = np.concatenate(b_list)
    keep = a != b
    a, b = a[keep], b[keep]
    return np.r_[a, b], np.r_[b, a]


def jones_plassmann(conn, seed=0):
    """Distance-1 colouring of the node-sharing graph (elements sharing a node get different colours).

    Each round selects the uncoloured elements whose random priority beats every uncoloured neighbour
    (an independent set) and gives each the smallest colour none of its coloured neighbours uses.
    """
    n_el = conn.shape[0]
    a, b = _node_sharing_pairs(conn)
    rng = np.random.default_rng(seed)
    prio = rng.permutation(n_el)
    colour = np.full(n_el, -1, np.int64)
    used = np.zeros(n_el, np.int64)              # bit c set: a neighbour has colour c
    while (colour < 0).any():
        un = colour < 0
        m = un[a] & un[b] & (prio[b] > prio[a])
        blocked = np.zeros(n_el, bool)
        blocked[a[m]] = True
        sel = np.flatnonzero(un & ~blocked)
        u = used[sel]
        c = np.zeros(len(sel), np.int64)
        while True:
            taken = (u >> c) & 1
            if not taken.any():
                break
            c += taken
        colour[sel] = c
        # tell the neighbours of the newly coloured elements
        nb = np.isin(b, sel) if len(sel) < n_el else np.ones(len(b), bool)
        np.bitwise_or.at(used, a[nb], np.int64(1) << colour[b[nb]])
    return colour



# quad4 meshes of at most this many elements use small-chunk plans (64-element chunks, DFEA_SMALL kernels): with
# 256-element chunks a 100x100 mesh keeps 40 of the 256 CUs busy.  Measured on MI355X (PCG iteration, FP64):
# 100x100 11.4 -> 8.7 us, 150x150 11.7 -> 10.6 us, 200x200 12.1 -> 12.9 us (slower), hence the bound
SMALL_QUAD4_MAX_ELEMS = 24000


def chunking_for(nen, n_el, chunking="auto"):
    """(chunk_elems or None for the element type's default, element-kind flag) of a mesh: "auto" picks
    small-chunk plans for quad4 meshes of at most SMALL_QUAD4_MAX_ELEMS elements; "small" / "large" force."""
    if chunking not in ("auto", "small", "large"):
        raise ValueError(f"chunking must be 'auto', 'small' or 'large', got {chunking!r}")
    small = nen == 4 and (chunking == "small" or (chunking == "auto" and n_el <= SMALL_QUAD4_MAX_ELEMS))
    return (_lib.load().dfea_chunk_elems(4 | _lib.SMALL), _lib.SMALL) if small else (None, 0)


# Tail-balanced plans.  A product or PCG launch of n chunk workgroups puts workgroup b on XCD b mod 8 and, on
# the GPU's 256 CUs, next to workgroups b +- 256 (MI355X: with 625 chunks CU j holds workgroups j, j + 256 and,
# for j < 113, j + 512).  When the chunks fill k whole rounds of 256 and part of one more, the CUs that carry an
# extra chunk set the kernel time (400 x 400 quad4: 113 CUs with three chunks, 143 with two).  A tail plan keeps
# k rounds of full chunks and cuts the remaining elements into chunks of half the size, one for each of up to
# 256 CUs, so the busiest CU carries k + 1/2 chunks' worth instead of k + 1.
CUS = 256
NXCD = 8


def tail_chunk_sizes(n_el, E, cus=CUS, nxcd=NXCD):
    """Element counts of the chunks of a tail-balanced plan of n_el elements in chunks of up to E, in chunk
    order; None where the plain plan (chunks of E) is kept: fewer than one or more than two whole rounds of
    `cus` chunks (beyond that the workgroups no longer all fit on the GPU at once), or a remainder of more
    than `cus` half chunks.  Chunk c of a launch of n workgroups runs on workgroup b = nxcd * pos + x where
    c = x * (n // nxcd) + min(x, n % nxcd) + pos (the kernels' XCD remap); the chunks whose workgroups come
    after the whole rounds (pos >= rounds * cus / nxcd) are the half chunks, and the last of them in element
    order takes what remains."""
    rounds = n_el // (cus * E)
    rest = n_el - rounds * cus * E
    half = E // 2
    if rounds < 1 or rounds > 2 or rest == 0 or rest > cus * half or half < 1 or (rounds * cus) % nxcd:
        return None
    n_half = -(-rest // half)
    n = rounds * cus + n_half
    base, r = divmod(n, nxcd)
    full_pos = rounds * cus // nxcd
    sizes = []
    for x in range(nxcd):
        nx = base + (1 if x < r else 0)
        sizes += [E] * min(nx, full_pos) + [half] * max(0, nx - full_pos)
    sizes = np.array(sizes, np.int64)
    halves = np.flatnonzero(sizes != E)
    sizes[halves[-1]] = rest - (n_half - 1) * half
    assert len(halves) == n_half and sizes.sum() == n_el
    return sizes


def _require(ok, what):
    if not ok:
        raise ValueError(f"invalid mesh plan: {what}")


class MeshPlan:
    """Integer tables of one mesh + Dirichlet mask (numpy arrays, see module docstring)."""

    def __init__(self, nodes, conn, fixed, chunk_elems=None, max_nodes=None, cell_ij=None, mode="auto",
                 balance=True):
        """balance: tail-balanced chunk sizes where ``tail_chunk_sizes`` gives them (partials plans of linear
        elements in their default chunk size); False keeps chunks of chunk_elems everywhere."""
        nodes = np.asarray(nodes, np.float64)
        conn = np.asarray(conn, np.int64)
        fixed = np.asarray(fixed, bool).reshape(-1, 2)
        n_el, nen = conn.shape
        N = nodes.shape[0]
        lib = _lib.load()
        E = chunk_elems or lib.dfea_chunk_elems(nen)
        MAXLN = max_nodes or lib.dfea_chunk_nodes()
        BS = lib.dfea_chunk_threads(nen)
        BS_halo = lib.dfea_chunk_threads(nen | _lib.HALO)
        if mode == "auto":
            mode = "partials"          # measured faster than halo plans on MI355X (fewer VGPRs, full waves)
        if mode not in ("halo", "partials"):
            raise ValueError(f"plan mode must be 'auto', 'halo' or 'partials', got {mode!r}")
        if E <= 0:
            raise ValueError(f"no chunk size for elements with {nen} nodes")
        if N >= (1 << 28) or n_el >= (1 << 28):
            raise ValueError("meshes with 2^28 or more nodes or elements are not supported (the kernels address "
                             "element planes and node arrays through 32-bit byte offsets)")
        self.n_el, self.nen, self.n_nodes = n_el, nen, N
        self.chunk_elems = E

        # 1. order
        if cell_ij is not None:                 # logical cell indices supplied by the mesh generator
            cij = np.asarray(cell_ij, np.int64)
            ix, iy, structured = cij[:, 0], cij[:, 1], True
        else:
            ix, iy, structured = cell_coordinates(nodes, conn)
        key = morton(ix, iy)
        order = np.argsort(key, kind="stable")
        # 2. colours
        if structured and _parity_colouring_valid(conn, ix, iy):
            colour = (ix % 2) * 2 + (iy % 2)
        else:
            colour = jones_plassmann(conn)
        if colour.max() >= _lib.MAXCOL:
            raise ValueError(f"mesh needs {colour.max() + 1} colours, more than {_lib.MAXCOL}")
        sizes = tail_chunk_sizes(n_el, E, CUS, NXCD) if (balance and mode == "partials" and nen <= 4 and
                                             E == lib.dfea_chunk_elems(nen)) else None
        self.tail = sizes is not None
        if sizes is None:
            n_chunks = (n_el + E - 1) // E
            sizes = np.minimum(E, n_el - np.arange(n_chunks) * E)
            chunk_of_pos = np.arange(n_el) // E                    # chunk of each Morton position
        else:
            # full chunks take the Morton order's whole tiles in chunk order, the half chunks the rest
            n_chunks = len(sizes)
            full = np.flatnonzero(sizes == E)
            halves = np.flatnonzero(sizes != E)
            chunk_of_pos = np.r_[np.repeat(full, E), np.repeat(halves, sizes[halves])]
        # within a chunk: sort by colour, keep Morton order inside a colour (internal order = chunk order)
        perm = order[np.lexsort((np.arange(n_el), colour[order], chunk_of_pos))]
        self.perm = perm.astype(np.int32)                          # internal e -> original element
        ce = conn[perm]                                            # (n_el, nen), internal order
        ch_e = np.repeat(np.arange(n_chunks), sizes)               # chunk of internal element
        col_sorted = colour[perm]

        bits = fixed[:, 0].astype(np.uint64) * np.uint64(_lib.FIXED_X) + \
            fixed[:, 1].astype(np.uint64) * np.uint64(_lib.FIXED_Y)
        self.nfix = (fixed[:, 0].astype(np.uint8) | (fixed[:, 1].astype(np.uint8) << 1))
        self.conn_int = np.ascontiguousarray(conn[perm].T).astype(np.int32)     # [nen][n_el]
        self.n_chunks = n_chunks
        self.structured = structured
        if mode == "halo" and BS_halo > 0 and self._build_halo(conn, perm, colour, E, BS_halo, MAXLN, bits):
            self.mode = "halo"
            return
        self.mode = "partials"
        self.hel = np.zeros(0, np.int32)
        self.hlelem = np.zeros((nen, 0), np.uint16)

        # colour ranges per chunk
        chunks = np.zeros((n_chunks, _lib.CHUNK_INTS), np.int32)
        ne = sizes
        e0 = np.cumsum(ne) - ne
        chunks[:, 0] = e0
        chunks[:, 1] = ne
        ncol = np.zeros(n_chunks, np.int32)
        cnt = np.zeros((n_chunks, _lib.MAXCOL), np.int64)
        np.add.at(cnt, (ch_e, col_sorted), 1)
        # compact the colours present in each chunk (empty colours are skipped)
        for c in range(n_chunks):
            present = cnt[c][cnt[c] > 0]
            ncol[c] = len(present)
            chunks[c, _lib.COL_OFF:_lib.COL_OFF + len(present) + 1] = np.r_[0, np.cumsum(present)]
        chunks[:, 6] = ncol

        # 3. local nodes: distinct (chunk, node) pairs, owned first
        pc = np.repeat(ch_e, nen)
        pn = ce.reshape(-1)
        pk = pc.astype(np.int64) * N + pn
        upk = np.unique(pk)
        uc, un = upk // N, upk % N
        touch = np.bincount(un, minlength=N)                       # chunks touching each node
        shared = touch[un] > 1
        k2 = uc * (2 * N) + shared * N + un                        # owned first inside each chunk
        o2 = np.argsort(k2, kind="stable")
        uc, un, shared, k2 = uc[o2], un[o2], shared[o2], k2[o2]
        k0 = np.searchsorted(uc, np.arange(n_chunks))
        nk = np.diff(np.r_[k0, len(uc)])
        if nk.max() > MAXLN:
            raise ValueError(f"a chunk touches {nk.max()} nodes, more than {MAXLN}")
        nown = np.bincount(uc[~shared], minlength=n_chunks)
        chunks[:, 2] = k0
        chunks[:, 3] = nk
        chunks[:, 4] = nown
        bits = fixed[:, 0].astype(np.uint64) * np.uint64(_lib.FIXED_X) + \
            fixed[:, 1].astype(np.uint64) * np.uint64(_lib.FIXED_Y)
        self.lnode = (un.astype(np.uint64) | bits[un]).astype(np.uint32)
        # local index of every element node
        key_e = pc.astype(np.int64) * (2 * N) + (touch[pn] > 1) * N + pn
        pos = np.searchsorted(k2, key_e)
        assert np.all(k2[pos] == key_e)
        loc = (pos - k0[pc]).reshape(n_el, nen)
        self.lelem = np.ascontiguousarray(loc.T).astype(np.uint16)      # [nen][n_el]

        # 4. shared nodes and partial slots
        sh_pos = np.flatnonzero(shared)                            # in (chunk, owned-first) order
        slot_of = np.full(len(uc), -1, np.int64)
        slot_of[sh_pos] = np.arange(len(sh_pos))
        pbase = np.zeros(n_chunks, np.int64)
        first_sh = k0 + nown
        has = nk > nown
        pbase[has] = slot_of[first_sh[has]]
        chunks[:, 5] = pbase
        so = np.lexsort((uc[sh_pos], un[sh_pos]))                  # by node, then chunk
        sh_nodes_all = un[sh_pos][so]
        chunk_major = slot_of[sh_pos][so]                            # node-major position -> chunk-major entry
        self.pslot = np.empty(len(chunk_major), np.int32)
        self.pslot[chunk_major] = np.arange(len(chunk_major), dtype=np.int32)
        snodes, sstart = np.unique(sh_nodes_all, return_index=True)
        self.sh_off = np.r_[sstart, len(sh_nodes_all)].astype(np.int32)
        self.sh_node = (snodes.astype(np.uint64) | bits[snodes]).astype(np.uint32)
        self.chunks = chunks
        self.conn_int = np.ascontiguousarray(ce.T).astype(np.int32)     # [nen][n_el]
        self.n_chunks = n_chunks
        self.n_lnodes = len(un)
        self.n_shared = len(snodes)
        self.n_part = len(sh_pos)
        self.structured = structured
        self.nfix = (fixed[:, 0].astype(np.uint8) | (fixed[:, 1].astype(np.uint8) << 1))

    def _build_halo(self, conn, perm, colour, E, BS, MAXLN, bits):
        """Halo-mode tables; returns False (and builds nothing) when a chunk would not fit the workgroup."""
        n_el, nen = conn.shape
        N = self.n_nodes
        n_chunks = self.n_chunks
        pos = np.empty(n_el, np.int64)
        pos[perm] = np.arange(n_el)                                # internal index of every element
        chunk_of = pos // E                                        # core chunk (by internal position)
        # owner of a node: chunk of its adjacent element latest in the internal order
        last = np.full(N, -1, np.int64)
        np.maximum.at(last, conn.reshape(-1), np.repeat(pos, nen))
        owner = last // E
        # (chunk, element) pairs: every chunk that owns a node of the element evaluates it
        pc = owner[conn].reshape(-1)
        pe = np.repeat(np.arange(n_el), nen)
        key = np.unique(pc * n_el + pe)
        hc, he = key // n_el, key % n_el
        halo = hc != chunk_of[he]
        hc, he = hc[halo], he[halo]
        ncol_g = int(colour.max()) + 1
        # halo entries sorted by (chunk, colour, internal index)
        o = np.lexsort((pos[he], colour[he], hc))
        hc, he = hc[o], he[o]
        h0 = np.searchsorted(hc, np.arange(n_chunks))
        nh = np.diff(np.r_[h0, len(hc)])
        ne = np.minimum(E, n_el - np.arange(n_chunks) * E)
        if (ne + nh).max() > BS:
            return False
        # local nodes: owned first, then the other nodes of the chunk's core and halo elements
        ce = conn[perm]
        core_pairs_c = np.repeat(np.arange(n_el) // E, nen)
        core_pairs_n = ce.reshape(-1)
        halo_pairs_c = np.repeat(hc, nen)
        halo_pairs_n = conn[he].reshape(-1)
        allc = np.r_[core_pairs_c, halo_pairs_c]
        alln = np.r_[core_pairs_n, halo_pairs_n]
        k = np.unique(allc * N + alln)
        uc, un = k // N, k % N
        own = owner[un] == uc
        k2 = uc * (2 * N) + (~own) * N + un                        # owned first inside each chunk
        o2 = np.argsort(k2, kind="stable")
        uc, un, own, k2 = uc[o2], un[o2], own[o2], k2[o2]
        k0 = np.searchsorted(uc, np.arange(n_chunks))
        nk = np.diff(np.r_[k0, len(uc)])
        if nk.max() > MAXLN:
            return False
        nown = np.bincount(uc[own], minlength=n_chunks)
        if nown.sum() != N:
            return False

        def local(chunks, nodes_):
            kk = chunks * (2 * N) + (owner[nodes_] != chunks) * N + nodes_
            ps = np.searchsorted(k2, kk)
            assert np.all(k2[ps] == kk)
            return ps - k0[chunks]

        self.lelem = np.ascontiguousarray(local(core_pairs_c, core_pairs_n).reshape(n_el, nen).T).astype(np.uint16)
        self.hlelem = np.ascontiguousarray(local(halo_pairs_c, halo_pairs_n).reshape(-1, nen).T).astype(np.uint16)
        self.hel = pos[he].astype(np.int32)                        # internal element ids
        self.lnode = (un.astype(np.uint64) | bits[un]).astype(np.uint32)
        chunks = np.zeros((n_chunks, _lib.CHUNK_INTS), np.int32)
        chunks[:, 0] = np.arange(n_chunks) * E
        chunks[:, 1] = ne
        chunks[:, 2] = k0
        chunks[:, 3] = nk
        chunks[:, 4] = nown
        chunks[:, 6] = ncol_g
        chunks[:, 7] = h0
        chunks[:, 8] = nh
        col_sorted = colour[perm]
        cc = np.zeros((n_chunks, ncol_g), np.int64)
        np.add.at(cc, (np.arange(n_el) // E, col_sorted), 1)
        hcc = np.zeros((n_chunks, ncol_g), np.int64)
        np.add.at(hcc, (hc, colour[he]), 1)
        chunks[:, _lib.COL_OFF:_lib.COL_OFF + ncol_g + 1] = np.c_[np.zeros(n_chunks, np.int64), np.cumsum(cc, 1)]
        chunks[:, _lib.HCOL_OFF:_lib.HCOL_OFF + ncol_g + 1] = np.c_[np.zeros(n_chunks, np.int64), np.cumsum(hcc, 1)]
        self.chunks = chunks
        self.n_lnodes = len(un)
        self.n_shared = 0
        self.n_part = 0
        self.pslot = np.zeros(0, np.int32)
        self.sh_node = np.zeros(0, np.uint32)
        self.sh_off = np.zeros(1, np.int32)
        return True

    @property
    def shmap(self):
        """Shared-node index of every node, -1 for nodes owned by one chunk."""
        m = np.full(self.n_nodes, -1, np.int32)
        m[(self.sh_node & _lib.NODE_MASK).astype(np.int64)] = np.arange(self.n_shared, dtype=np.int32)
        return m

    @property
    def shslot(self):
        """[n_nodes][2] int32: first partial slot and number of partials of every node ({0, 0} if owned)."""
        out = np.zeros((self.n_nodes, 2), np.int32)
        nodes = (self.sh_node & _lib.NODE_MASK).astype(np.int64)
        out[nodes, 0] = self.sh_off[:-1]
        out[nodes, 1] = np.diff(self.sh_off)
        return out

    def inverse_map(self):
        """Node-to-element map of the product's node sums (partials plans): ``inv_off`` [n_lnodes + 1] int32 and
        ``inv_slot`` [n_el * nen] uint16.  The element vectors of local node entry i (global entry index
        k0 + k) are the chunk-local slots inv_slot[inv_off[i] .. inv_off[i+1]), slot = e_local * nen + a
        (element e_local of the chunk, its node a), sorted by slot, i.e. in element order, which inside a chunk
        is colour order: the sums come out in the same order as colour-by-colour accumulation.  A chunk's
        slots occupy inv_slot[e0 * nen .. (e0 + ne) * nen)."""
        ch = self.chunks
        nen = self.nen
        chunk_of_e = np.repeat(np.arange(self.n_chunks), ch[:, 1])
        e_local = np.arange(self.n_el) - ch[chunk_of_e, 0]
        entry = (ch[chunk_of_e, 2][None, :] + self.lelem.astype(np.int64)).T.reshape(-1)   # (e, a) -> entry
        slot = (e_local[:, None] * nen + np.arange(nen)[None, :]).reshape(-1)
        order = np.lexsort((slot, entry))
        inv_slot = slot[order].astype(np.uint16)
        inv_off = np.searchsorted(entry[order], np.arange(self.n_lnodes + 1)).astype(np.int32)
        return inv_off, inv_slot

    def device_tables(self, inv_off=None):
        """The local-node tables as the kernels read them: dict(chunks, lnode, pslot, prange, inv_off, lstride).

        Partials plans are laid out at a fixed stride: local-node entry k of chunk c is c * lstride + k
        (lstride = the largest chunk node count rounded up to 32), for the node table, the partial slot and
        slot range of shared entries, and the inverse-map offsets, so a workgroup addresses its node tables
        from its chunk index alone, without waiting for its chunk record.  The chunk records point into that
        layout (k0 = c * lstride, pbase = c * lstride + nown), so kernels that read them see the same entries.
        Padding entries hold the chunk's last node (node table), slot 0 and range {0, 0} (owned and padding
        entries), and the chunk's end offset (inverse map).  Halo plans keep the compact tables (lstride 0).
        inv_off: this plan's ``inverse_map()[0]`` (partials plans)."""
        if self.mode != "partials":
            return dict(chunks=self.chunks, lnode=self.lnode, pslot=self.pslot, prange=self.prange,
                        inv_off=inv_off, lstride=0)
        ch = self.chunks
        nc = self.n_chunks
        k0, nk, nown = ch[:, 2].astype(np.int64), ch[:, 3].astype(np.int64), ch[:, 4].astype(np.int64)
        ls = int(-(-int(nk.max()) // 32) * 32)
        _require(nc * ls < (1 << 31), "padded node tables beyond 2^31 entries")
        k = np.arange(ls)[None, :]
        src = k0[:, None] + np.minimum(k, nk[:, None] - 1)        # compact entry of every padded entry
        lnode = self.lnode[src.reshape(-1)]
        live_sh = (k >= nown[:, None]) & (k < nk[:, None])         # shared entries of each chunk
        pslot = np.zeros((nc, ls), np.int32)
        prange = np.zeros((nc, ls, 2), np.int32)
        if self.n_part:
            ent = (ch[:, 5].astype(np.int64)[:, None] + (k - nown[:, None]))[live_sh]
            pslot[live_sh] = self.pslot[ent]
            prange[live_sh] = self.prange[ent]
        dch = ch.copy()
        dch[:, 2] = np.arange(nc) * ls
        dch[:, 5] = np.arange(nc) * ls + nown
        out = dict(chunks=dch, lnode=lnode, pslot=pslot.reshape(-1), prange=prange.reshape(-1, 2), lstride=ls)
        if inv_off is not None:
            inv_off = np.asarray(inv_off, np.int64)
            _require(len(inv_off) == self.n_lnodes + 1, "inverse map size")
            end = np.r_[k0[1:], self.n_lnodes]                    # compact entry after each chunk's last node
            inv = inv_off[np.minimum(k0[:, None] + k, end[:, None])]   # entries past nk: the chunk's end offset
            out["inv_off"] = np.r_[inv.reshape(-1), inv_off[-1]].astype(np.int32)
        return out

    @property
    def prange(self):
        """[n_part][2] int32: first slot and number of slots of the shared node of every partials entry
        (chunk order, like pslot)."""
        s = np.searchsorted(self.sh_off, self.pslot, side="right") - 1      # shared index of each entry
        return np.stack([self.sh_off[s], self.sh_off[s + 1] - self.sh_off[s]], axis=1).astype(np.int32)

    def validate(self, max_nodes=None, inv=None):
        """Host-side bounds checks of every index the kernels follow (run before any launch); raises ValueError
        on the first violation.  max_nodes: the kernels' local-node capacity per chunk (LDS), checked when
        given; inv: the (inv_off, inv_slot) tables of a partials plan, checked when given.
        Returns the largest number of threads a chunk needs (core + halo elements)."""
        ch = self.chunks
        nen = self.nen
        _require(ch[:, 0].min() >= 0 and (ch[:, 0] + ch[:, 1]).max() <= self.n_el, "chunk element range")
        _require((ch[:, 2] + ch[:, 3]).max() <= self.n_lnodes and (ch[:, 4] <= ch[:, 3]).all(), "chunk node range")
        _require(ch[:, 3].min() >= 1, "chunk without nodes")
        if max_nodes is not None:
            _require(int(ch[:, 3].max()) <= max_nodes,
                     f"a chunk touches {int(ch[:, 3].max())} nodes, more than the kernels' {max_nodes}")
        if self.mode == "partials":
            _require((ch[:, 5] + ch[:, 3] - ch[:, 4]).max() <= max(self.n_part, 0), "chunk partials range")
            _require((ch[:, 8] == 0).all(), "halo elements in a partials plan")
        else:
            _require(self.n_part == 0 and ch[:, 4].sum() == self.n_nodes, "halo plan ownership")
            _require((ch[:, 7] + ch[:, 8]).max() <= len(self.hel), "halo range")
            if len(self.hel):
                _require(self.hel.min() >= 0 and self.hel.max() < self.n_el, "halo element id")
                ch_of_h = np.repeat(np.arange(self.n_chunks), ch[:, 8])
                _require((self.hlelem.astype(np.int64) < ch[ch_of_h, 3][None, :]).all(), "halo local node index")
        # every element's local indices fall inside its chunk's node range
        chunk_of_e = np.repeat(np.arange(self.n_chunks), ch[:, 1])
        _require((self.lelem.astype(np.int64) < ch[chunk_of_e, 3][None, :]).all(), "element local node index")
        # colour ranges tile the core and halo threads
        nc = ch[:, 6]
        _require(((nc >= 1) & (nc <= _lib.MAXCOL)).all(), "colour count")
        rows = np.arange(self.n_chunks)
        _require((ch[:, _lib.COL_OFF] == 0).all() and (ch[rows, _lib.COL_OFF + nc] == ch[:, 1]).all(), "colour ranges")
        _require((ch[:, _lib.HCOL_OFF] == 0).all() and (ch[rows, _lib.HCOL_OFF + nc] == ch[:, 8]).all(),
                 "halo colour ranges")
        for c in range(_lib.MAXCOL):      # colour ranges non-decreasing up to the colour count
            live = c < nc
            _require((ch[live, _lib.COL_OFF + c] <= ch[live, _lib.COL_OFF + c + 1]).all(), "colour order")
        _require((self.lnode & _lib.NODE_MASK).max() < self.n_nodes, "local node id")
        if self.n_shared:
            _require(self.sh_off[0] == 0 and self.sh_off[-1] == len(self.pslot) == self.n_part, "shared slot ranges")
            _require((np.diff(self.sh_off) >= 1).all(), "shared node without slots")
            _require(np.array_equal(np.sort(self.pslot), np.arange(self.n_part)), "partial slots")
            _require((self.sh_node & _lib.NODE_MASK).max() < self.n_nodes, "shared node id")
            _require(len(np.unique(self.sh_node & _lib.NODE_MASK)) == self.n_shared, "shared node repeated")
            _require(np.array_equal(self.sh_node >> 30, self.nfix[self.sh_node & _lib.NODE_MASK]), "shared fixed bits")
            pr = self.prange
            _require(pr[:, 0].min() >= 0 and (pr[:, 1] >= 1).all() and (pr[:, 0] + pr[:, 1]).max() <= self.n_part,
                     "partials entry slot range")
            _require(((self.pslot >= pr[:, 0]) & (self.pslot < pr[:, 0] + pr[:, 1])).all(),
                     "a chunk's slot outside its node's range")
        if inv is not None:
            inv_off, inv_slot = inv
            _require(len(inv_off) == self.n_lnodes + 1 and len(inv_slot) == self.n_el * nen, "inverse map size")
            _require(inv_off[0] == 0 and (np.diff(inv_off) >= 0).all(), "inverse map offsets not monotone")
            _require((inv_off[ch[:, 2]] == ch[:, 0] * nen).all() and
                     (inv_off[ch[:, 2] + ch[:, 3]] == (ch[:, 0] + ch[:, 1]) * nen).all(),
                     "inverse map offsets outside their chunk")
            ent_chunk = np.repeat(np.arange(self.n_chunks), ch[:, 1] * nen)
            _require((inv_slot.astype(np.int64) < ch[ent_chunk, 1] * nen).all(), "inverse map slot")
        _require(self.conn_int.min() >= 0 and self.conn_int.max() < self.n_nodes, "connectivity")
        return int((ch[:, 1] + ch[:, 8]).max())

    def stats(self):
        return dict(mode=self.mode, n_chunks=self.n_chunks, n_lnodes=self.n_lnodes, n_shared=self.n_shared,
                    n_part=self.n_part, n_halo=len(self.hel), max_chunk_nodes=int(self.chunks[:, 3].max()),
                    max_colours=int(self.chunks[:, 6].max()), max_threads=int((self.chunks[:, 1] + self.chunks[:, 8]).max()))
