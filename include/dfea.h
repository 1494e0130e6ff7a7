/*
 * dfea.h -- C ABI of libdfea_hip.so, the MI355X (gfx950) matrix-free Newton-Krylov hot path of
 * Differentiable-FEA (plane-strain hyperelastic material library; tri3, quad4, tri6, quad8 and quad9 elements).
 *
 * Conventions
 *   - Every entry point returns 0 on success, DFEA_EINVAL (-1) for an invalid argument, or
 *     -(1000 + hipError_t) when a HIP call failed.  dfea_last_error() gives a message for the last failure.
 *   - All work is enqueued on the caller's stream (hipStream_t passed as void*); no entry point
 *     synchronises the stream (callers read the scalar block, e.g. the PCG status, when they choose).
 *   - The caller allocates and owns every buffer.  Descriptors (dfea_desc) hold device pointers into those
 *     buffers; a descriptor table in device memory describes a batch of independent solves.
 *   - Element kind (`nen` arguments): the nodes per element, 3 (tri3, 1 Gauss point), 4 (quad4, 2x2),
 *     6 (tri6, 6-point degree-4 rule), 8 (quad8, 3x3) or 9 (quad9, 3x3), with nq of the descriptors equal to
 *     that rule's size; 4 | DFEA_HALO for quad4 halo plans, 4 | DFEA_SMALL for small-chunk quad4 plans,
 *     3 | DFEA_TAIL or 4 | DFEA_TAIL for tail-balanced plans, 6, 8 or 9 | DFEA_WIDE for small quadratic-element
 *     meshes (PCG workgroups); | DFEA_MATLIB when a descriptor's material is not neo-Hookean.
 *   - Work items: an int32 pair {solve index, block index} per workgroup; for a batch, the concatenation
 *     over solves.  Any item table may be NULL for one solve: solve 0 (the table's first descriptor),
 *     blocks 0 .. n_items-1, which saves the workgroups one dependent load.
 *
 * Data layout in HBM (internal element order = chunk order, see dfea_chunk)
 *   Per-element arrays are planes of n_cols = n_el + n_halo columns: column e < n_el is element e, column
 *   n_el + h holds a copy for halo entry h (halo plans), so halo threads read coalesced like core threads.
 *   dN[(q*nen + a)*2 + j][col] shape-function gradients dN_a/dX_j at Gauss point q
 *   w[q][col]                  Gauss weight * det(dX/dxi)
 *   theta[col]                 design factor
 *   C[k*nq + q][col]           the 10 unique symmetric tangent moduli, k over
 *                              (00,01,02,03,11,12,13,22,23,33) of the flat index m = 2i+j,
 *                              already multiplied by w (`src/difffea/operators.py:316-342`)
 *   vectors                    node-major dofs 2a+i, as the reference (`boundary_conditions.py:171-179`)
 */
#ifndef DFEA_H
#define DFEA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DFEA_EINVAL (-1)
#define DFEA_MAXCOL 16         /* element colours per chunk */
#define DFEA_NOALLOC_SCRATCH_BYTES (8 << 20)   /* static device scratch of dfea_apply_noalloc, per GPU */
#define DFEA_FIXED_X 0x40000000u
#define DFEA_FIXED_Y 0x80000000u
#define DFEA_NODE_MASK 0x3fffffffu
#define DFEA_HALO 0x100        /* element-kind flag: nen | DFEA_HALO selects the halo-plan kernels */
#define DFEA_SMALL 0x400       /* element-kind flag: 4 | DFEA_SMALL selects the quad4 kernels of small-chunk plans
                                  (64-element chunks; the product and PCG give each element 4 threads) */
#define DFEA_TAIL 0x800        /* element-kind flag: 3 | DFEA_TAIL or 4 | DFEA_TAIL for tail-balanced partials plans
                                  (chunks of dfea_chunk_elems elements and, past whole rounds of 256 chunks,
                                  chunks of half that size): the product and PCG read each chunk's element
                                  range from its record */
#define DFEA_WIDE 0x1000       /* element-kind flag: 6, 8 or 9 | DFEA_WIDE for quadratic-element meshes of at most 256
                                  chunks (one workgroup per CU or fewer): the PCG kernel gives each element more
                                  Gauss-point groups (192 threads per chunk in FP64, 256 in FP32, instead of 128 and
                                  192), whatever the size of the launch the mesh's solve runs in */
#define DFEA_MATLIB 0x200      /* element-kind flag: the descriptors may use any DFEA_MAT_* material; without it
                                  the material kernels compile only the neo-Hookean closed forms (fewer
                                  registers) and give NaN for a descriptor of another material */

enum { DFEA_FP64 = 0, DFEA_FP32 = 1 };
/* dfea_desc.material: the energies of `src/difffea/materials.py:65-162`, parameters in that file's order
 * (neo-Hookean [mu, lmbda]; Mooney-Rivlin [c10, c01, kappa]; Saint Venant-Kirchhoff [mu, lmbda]; Demiray
 * [mu, gamma, lmbda]; fibre-reinforced HGO [mu, lmbda, k1, k2, beta_deg, d, second_family]) */
enum {
    DFEA_MAT_NEO_HOOKEAN = 0, DFEA_MAT_MOONEY_RIVLIN = 1, DFEA_MAT_SAINT_VENANT = 2, DFEA_MAT_DEMIRAY = 3,
    DFEA_MAT_FIBER_REINFORCED = 4
};
#define DFEA_MAT_NP 8          /* parameter slots of dfea_desc.mat_p */
/* dfea_desc.precond: point Jacobi (`operators.py:477-570`) or nodal block Jacobi (`operators.py:588-638`) */
enum { DFEA_PRECOND_JACOBI = 0, DFEA_PRECOND_BLOCK = 1 };

/* vector slots of a descriptor.  The PCG state x, r, w = A M r, p, s = A p lives in two buffers each (the
 * slot and its "2" twin): iteration k reads the buffers of parity k mod 2 and writes the others, so no
 * workgroup overwrites a value another workgroup still has to read.  The PCG solution ends in vec[DFEA_X]. */
enum { DFEA_X = 0, DFEA_R = 1, DFEA_W = 2, DFEA_P = 3, DFEA_S = 4, DFEA_B = 5, DFEA_U = 6, DFEA_RES = 7,
       DFEA_UACC = 8, DFEA_UD = 9, DFEA_X2 = 10, DFEA_R2 = 11, DFEA_W2 = 12, DFEA_P2 = 13, DFEA_S2 = 14,
       DFEA_NVEC = 15 };

/* scalar block of a solve (doubles).  Slots 0..10 are the PCG state read by the caller (RZ = r.M r,
 * PAP = p.A p and ALPHA / BETA of the last update, RR = r.r of the last tested residual, K = updates done);
 * the PCG kernels themselves carry their state in two parity sets of 4 doubles at DFEA_SC_S0 / DFEA_SC_S1
 * (status, k, r.M r, alpha), each written by the kernels of one parity and read by those of the other. */
enum { DFEA_SC_RZ = 0, DFEA_SC_PAP = 1, DFEA_SC_ALPHA = 2, DFEA_SC_BETA = 3, DFEA_SC_RR = 4, DFEA_SC_BB = 5,
       DFEA_SC_RTOL = 6, DFEA_SC_STATUS = 7, DFEA_SC_K = 8, DFEA_SC_MAXIT = 9, DFEA_SC_ITERS = 10,
       DFEA_SC_NORM = 11, DFEA_SC_LF = 12, DFEA_SC_STEP = 13, DFEA_SC_S0 = 16, DFEA_SC_S1 = 20, DFEA_NSC = 24 };

/* PCG status values stored in sc[DFEA_SC_STATUS] */
enum { DFEA_PCG_RUNNING = 0, DFEA_PCG_CONVERGED = 1, DFEA_PCG_CURVATURE = 2, DFEA_PCG_MAXITER = 3 };

/* A chunk: the elements handled by one workgroup (dfea_chunk_elems / dfea_chunk_nodes bound its size).
 * Two plan modes:
 *   partials  nodes used by one chunk are owned by it; nodes used by several chunks get one partial per
 *             chunk (pslot), summed by the *_fin kernels in chunk order;
 *   halo      every node is owned by one chunk, and the chunk also evaluates the "halo" elements of other
 *             chunks that touch its owned nodes, so each owned node is completed inside the workgroup (no
 *             shared nodes, no second kernel).  Halo threads only add into LDS; per-element outputs
 *             (moduli, geometry, adjoint gradient) are written by the element's core chunk. */
typedef struct dfea_chunk {
    int32_t e0, ne;        /* core elements [e0, e0+ne) (internal order) = threads 0..ne-1 */
    int32_t k0, nk;        /* local node table range [k0, k0+nk) in lnode */
    int32_t nown;          /* local nodes k0..k0+nown-1 are owned (final values written by this chunk) */
    int32_t pbase;         /* partials mode: first pslot entry of the chunk's shared local nodes */
    int32_t ncol;          /* element colours */
    int32_t h0, nh;        /* halo mode: halo entries [h0, h0+nh) of hel / hlelem = threads ne..ne+nh-1 */
    int32_t pad[3];
    int32_t col[DFEA_MAXCOL + 1];   /* colour c = core threads col[c]..col[c+1]-1 */
    int32_t hcol[DFEA_MAXCOL + 1];  /* colour c = halo threads ne+hcol[c]..ne+hcol[c+1]-1 */
    int32_t pad2[18];      /* 256 bytes */
} dfea_chunk;

/* Everything the kernels need about one solve (device pointers). */
typedef struct dfea_desc {
    const dfea_chunk* chunks;   /* [n_chunks] */
    const uint32_t* lnode;      /* [n_lnodes] node id | DFEA_FIXED_X | DFEA_FIXED_Y (see lstride) */
    const uint16_t* lelem;      /* [nen][n_el] local index (within the chunk) of element node a */
    const uint32_t* sh_node;    /* [n_shared] node id | fixed bits */
    const int32_t* sh_off;      /* [n_shared+1] */
    const int32_t* pslot;       /* partial slot of each (chunk, shared node) entry: entry pbase + (k - nown) of a
                                   chunk's shared local node k ([n_lnodes] at the lstride layout, else [n_part]);
                                   the partials of shared node i are slots sh_off[i]..sh_off[i+1]-1 */
    const int32_t* hel;         /* [n_halo] halo mode: internal element id of each halo entry */
    const uint16_t* hlelem;     /* [nen][n_halo] halo mode: local node index of halo entry h, node a */
    const int32_t* prange;      /* {first slot, number of slots} of the shared node of each (chunk, shared node)
                                   entry, indexed like pslot (entry tables load without the node id) */
    const int32_t* inv_off;     /* [n_lnodes + 1] partials plans: the element vectors of local node entry i are the
                                   chunk-local slots inv_slot[inv_off[i] .. inv_off[i+1]) */
    const uint16_t* inv_slot;   /* [n_el * nen] slot e_local * nen + a, sorted by slot per node entry (element =
                                   colour order); chunk c's slots occupy [e0 * nen, (e0 + ne) * nen) */
    const int32_t* conn;        /* [nen][n_el] node ids, internal element order */
    const int32_t* perm;        /* [n_el] original element id of internal element e */
    const double* nodes;        /* [N][2] reference coordinates */
    const double* dNdxi;        /* [nq][nen][2] reference-element gradients at the Gauss points */
    const double* gw;           /* [nq] Gauss weights */
    const uint8_t* nfix;        /* [N] bit 0: x fixed, bit 1: y fixed */
    double* dN64;
    float* dN32;
    double* w;
    double* theta;              /* [n_cols] internal order (+ halo copies) */
    double* C64;
    float* C32;
    double* diag64;             /* Jacobi: [2N] diag(A); block Jacobi: [3N], the inverse nodal blocks {m00, m11}
                                   per node, then m01 per node */
    float* diag32;
    double* f;                  /* [2N] external force at load factor 1 */
    double* part;               /* shared-node partials, 4 n_part doubles: the product and the Newton-step kernels
                                   use n_part pairs of their precision from the start (block-Jacobi moduli also the
                                   second half); the PCG's two parities are FP64 pairs [q n_part, (q + 1) n_part)
                                   or FP32 pairs [(2 + q) n_part, (3 + q) n_part), so an FP64 residual of the same
                                   solve may run between two rounds of PCG kernels */
    double* red;                /* reduction partials, dfea_red_size() doubles */
    double* sc;                 /* [DFEA_NSC] */
    void* vec[DFEA_NVEC];       /* PCG / Newton vectors; x,r,z,p,Ap,b in the solve's precision; u (evaluation
                                   point), R, u_acc (accepted state), u_D (prescribed values) FP64 */
    double mat_p[DFEA_MAT_NP];  /* nominal material parameters; entry k is multiplied by theta_e when bit k of
                                   mat_scal is set (`materials.py:239-246`) */
    int32_t n_chunks, n_lnodes, n_shared, n_part;
    int32_t n_el, n_nodes, nen, nq;
    int32_t precision, n_finblk, n_vblk, n_halo;
    int32_t n_cols;            /* plane length n_el + n_halo */
    int32_t precond;           /* DFEA_PRECOND_JACOBI or DFEA_PRECOND_BLOCK */
    int32_t material;          /* DFEA_MAT_* */
    int32_t mat_scal;          /* bit k: mat_p[k] is theta-scaled */
    int32_t lstride;           /* partials plans: the local-node tables (lnode, pslot, prange, inv_off) hold chunk c's
                                  entries at c * lstride + k, k < lstride (k >= nk: padding -- the chunk's last node,
                                  slot 0 / range {0, 0}, the chunk's end offset), so the product and PCG kernels
                                  address them from the chunk index alone; the chunk records agree (k0 = c * lstride,
                                  pbase = c * lstride + nown).  0 for halo plans (compact tables) */
} dfea_desc;

/* ---- library ---------------------------------------------------------------------------------------- */
int dfea_version(void);                                 /* 100*major + minor */
const char* dfea_last_error(void);
int dfea_desc_size(void);                               /* sizeof(dfea_desc), for binding checks */
int dfea_desc_offset(int field);                        /* byte offset of field #field (declaration order) */
int dfea_red_size(int n_chunks, int n_finblk, int n_vblk);   /* doubles needed for dfea_desc.red */
int dfea_vblk(int n_dof);                               /* dof blocks of the vector kernels */
int dfea_finblk(int n_shared);                          /* blocks of the shared-node kernels */
int dfea_chunk_elems(int nen);                          /* maximum elements per chunk for this element */
int dfea_chunk_nodes(void);                             /* maximum local nodes per chunk */
int dfea_chunk_threads(int kind);                       /* workgroup size of an element kind */

/* ---- once per mesh ---------------------------------------------------------------------------------- */
/* Geometry cache dN_dX, w_detJ (dN32 too when non-null).
 * Replaces `src/difffea/operators.py:85-128` (precompute_geometry / get_geometry). */
int dfea_geometry(const dfea_desc* table, const int32_t* items, int n_items, int nen, void* stream);

/* ---- once per Newton step --------------------------------------------------------------------------- */
/* Tangent moduli C (10 unique entries, weight folded in) at state u, stored FP64 (precision FP64) or FP32,
 * fused with the preconditioner of desc.precond into diag64/diag32: the Jacobi diagonal (clamp 1e-6 on free
 * dofs, 1 on fixed dofs), or the inverses of the nodal 2x2 blocks (fixed dofs: identity rows and columns;
 * + 1e-12 I), summed in FP64 and inverted in closed form.  The shared-node partials of the block's
 * off-diagonal entry use the second half of desc.part.  u == NULL uses vec[DFEA_U].
 * Replaces `operators.py:316-342` (_qp_setup), `operators.py:477-570` (compute_jacobi_diagonal) and
 * `operators.py:588-638` (make_block_jacobi_preconditioner).
 * items: chunk items; fin_items: shared-node items (dfea_finblk blocks per solve). */
int dfea_moduli(const dfea_desc* table, const int32_t* items, int n_items, const int32_t* fin_items,
                int n_fin_items, const double* u, int nen, int precision, void* stream);

/* ---- residual --------------------------------------------------------------------------------------- */
/* R = m * (R_int(u) - lf * f) with the analytic Piola stress; lf from sc[DFEA_SC_LF] when lf < 0.
 * u/R NULL -> vec[DFEA_U] / vec[DFEA_RES].  Replaces `operators.py:131-220` (global_residual). */
int dfea_residual(const dfea_desc* table, const int32_t* items, int n_items, const int32_t* fin_items,
                  int n_fin_items, const double* u, double* R, double lf, int nen, void* stream);

/* ---- post-processing ---------------------------------------------------------------------------------- */
/* Support reactions: out[2 s], out[2 s + 1] = the sums over the fixed dofs of the internal force R_int(u)
 * (all-ones mask, no external load) for every solve s of solve_items (NULL: solve 0, n_solves = 1).  The chunk
 * sums use the first 2 n_chunks doubles of desc.red (the PCG's partials): not while a PCG of the same solve
 * runs.  u NULL -> vec[DFEA_U].  Replaces `src/difffea/postprocess.py:121-137` (support_reactions). */
int dfea_reactions(const dfea_desc* table, const int32_t* items, int n_items, const int32_t* solve_items,
                   int n_solves, const double* u, double* out, int nen, void* stream);
/* Element and nodal fields of u (plane strain, the descriptor's material), items of one solve.
 * eout [7][n_el], original element order: area, then the area-weighted means of J, Cauchy sxx, syy, sxy, the
 * von Mises stress of the mean stress (with the plane-strain szz) and the energy density psi.
 * nout [6][N]: J, sxx, syy, sxy, mises, psi summed over the elements at each node and divided by their number
 * (nodes that no element touches are not written).  work: 8 * n_part doubles for the shared-node partials
 * (NULL when the plan has no shared nodes).  u NULL -> vec[DFEA_U].
 * Replaces `src/difffea/postprocess.py:24-118` (gauss_point_state, element_state, compute_fields). */
int dfea_fields(const dfea_desc* table, const int32_t* items, int n_items, const int32_t* fin_items,
                int n_fin_items, const double* u, double* eout, double* nout, double* work, int nen, void* stream);

/* ---- tangent product -------------------------------------------------------------------------------- */
/* y = m * J(u)(m v) + (1-m) v from the stored moduli (FP64 or FP32 per descriptor precision).
 * v/y NULL -> vec[DFEA_P] / vec[DFEA_W].  Deterministic (no atomics).
 * Replaces the closure of `operators.py:345-412` (make_A_operator_qp / _qp_ew). */
int dfea_apply(const dfea_desc* table, const int32_t* items, int n_items, const int32_t* fin_items,
               int n_fin_items, const void* v, void* y, int nen, int precision, void* stream);

/* The product of one solve (the table's first descriptor) given also the caller's host copy of that descriptor:
 * the shared-node finish takes its node tables and partials as kernel arguments (they arrive preloaded in
 * registers), so its loads do not wait for the descriptor.  host_desc must equal the table's first descriptor;
 * v and y are required.  Same result, bit for bit, as dfea_apply.  Replaces the closures of
 * `make_A_operator_qp` / `_qp_ew` (`src/difffea/operators.py:345-412`) for a single solve. */
int dfea_apply_host(const dfea_desc* host_desc, const dfea_desc* table, int n_chunks, int n_fin_blocks, const void* v,
                    void* y, int nen, int precision, void* stream);

/* The same product for ONE solve without any device allocation (no counterpart in the reference: it is for
 * callers that cannot allocate -- code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The shared-node partials that dfea_apply keeps in the descriptor's part buffer go instead to
 * a static device array of DFEA_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends; its address is
 * looked up once per device, with that device current).  table: one descriptor (the solve); n_chunks and
 * n_fin_blocks: the work-item counts dfea_apply takes for that solve alone (items and fin_items NULL);
 * n_part: the plan's number of shared-node partials (dfea_desc.n_part), at most
 * DFEA_NOALLOC_SCRATCH_BYTES / 16 in FP64 and / 8 in FP32, else DFEA_EINVAL and nothing is launched.
 * v and y must not be NULL.  The current device must be the stream's.  No hipMalloc, hipFree or
 * allocating library call.  Every partial is written before it is read, so nothing needs clearing
 * between calls; calls that overlap on one GPU (other streams, other host threads) share the array and
 * race. */
int dfea_apply_noalloc(const dfea_desc* table, int n_chunks, int n_fin_blocks, int n_part, const void* v, void* y,
                       int nen, int precision, void* stream);

/* ---- deterministic reductions ----------------------------------------------------------------------- */
/* sc[DFEA_SC_NORM] = sum(x*y) over the solve's dofs (x==y gives the squared norm); FP64 vectors,
 * x/y NULL -> vec[DFEA_RES].  items: dof-block items (dfea_vblk per solve); solve_items: {s, 0} per solve.
 * Replaces torch.dot / torch.norm at `newton_krylov.py:56,65,78,116,142`. */
int dfea_dot(const dfea_desc* table, const int32_t* items, int n_items, const int32_t* solve_items,
             int n_solves, const double* x, const double* y, void* stream);

/* ---- Newton helpers ----------------------------------------------------------------------------------- */
/* mode 0: out = where(fixed, lf*uD, u)   (the load-level lift, `newton_krylov.py:139`);
 * mode 1: out = u + alpha * dx           (the line-search trial, `newton_krylov.py:114`), dx in the solve's
 *                                         precision;
 * mode 2: out = u                        (accept a trial state).
 * NULL pointers select the descriptor's vectors: u -> vec[DFEA_UACC], dx -> vec[DFEA_X], uD -> vec[DFEA_UD],
 * out -> vec[DFEA_U] (mode 2: u -> vec[DFEA_U], out -> vec[DFEA_UACC]).  alpha / lf come from the per-solve
 * doubles sarg[solve]. */
int dfea_axpy(const dfea_desc* table, const int32_t* items, int n_items, int mode, const double* u,
              const void* dx, const double* uD, const double* sarg, double* out, int precision, void* stream);

/* ---- PCG ---------------------------------------------------------------------------------------------- */
/* Start a PCG solve: x = 0, r = b, p = s = 0 and the scalar block (rtol, max_iter are read from sc).
 * b: vec[DFEA_B] (solve precision), or b = -R from vec[DFEA_RES] with neg_res.  A zero b finishes at once
 * (converged, 0 iterations, x = 0).  Replaces `newton_krylov.py:51-60`. */
int dfea_pcg_init(const dfea_desc* table, const int32_t* vitems, int n_vitems, const int32_t* solve_items,
                  int n_solves, int neg_res, int precision, void* stream);
/* The same start for one solve (the table's first descriptor, NULL item tables over n_vitems vector blocks),
 * with the relative tolerance and the cap passed by value: the set-up kernel writes them into the scalar block,
 * so the caller copies nothing to the device first.  Replaces `newton_krylov.py:51-60` in a single solve's
 * Newton loop (`newton_krylov.py:147-148`). */
int dfea_pcg_start(const dfea_desc* table, int n_vitems, double rtol, int max_iter, int neg_res, int precision,
                   void* stream);
/* PCG kernels for every running solve of the batch, one kernel per iteration (per chunk: the scalar step
 * -- convergence of the last residual, the iteration cap, curvature, alpha and beta -- from the previous
 * kernel's partials, evaluated identically by every workgroup; the update of the chunk's nodes; the next
 * product A M r in LDS; partials of r.M r, (A M r).M r and r.r).  The recurrences are the Chronopoulos-Gear
 * arrangement of Jacobi-PCG: the iterates of `newton_krylov.py:62-88` in exact arithmetic, with its tests
 * in its order (non-positive curvature at the first iteration answers z = M b).
 * Kernels run in pairs of parities 0, 1: n_iter is rounded up to even, and dfea_pcg_init may only be
 * called between two calls.  The first kernel after dfea_pcg_init forms A M b; each further kernel is one
 * iteration, whose test on the residual it produced is made by the next kernel.  Callers launch while
 * sc[DFEA_SC_STATUS] reads DFEA_PCG_RUNNING; finished solves skip their work.  items: chunk items.
 * Between two calls the solve's residual, dot products and axpy operations may run on its descriptor, with
 * outputs other than the PCG vectors (they touch neither the partials the next kernel reads nor the scalars
 * it uses); the moduli may not (the PCG reads them every iteration). */
int dfea_pcg_iterate(const dfea_desc* table, const int32_t* items, int n_items, int n_iter, int nen, int precision,
                     void* stream);

/* Capture n_iter kernels of dfea_pcg_iterate (same arguments) into an executable HIP graph; replay it
 * with dfea_graph_launch on any stream.  The tables and buffers the descriptors point to must outlive the
 * graph.  No memset nodes are recorded.  Graph capture of the iteration loop replaces the reference's
 * per-iteration Python loop with two host syncs (`newton_krylov.py:62-86`). */
int dfea_pcg_graph_create(const dfea_desc* table, const int32_t* items, int n_items, int n_iter, int nen,
                          int precision, void** graph);
int dfea_graph_launch(void* graph, void* stream);
int dfea_graph_destroy(void* graph);

/* ---- adjoint -------------------------------------------------------------------------------------------- */
/* g[e] = -sum_{a,i} (m*lam)[2 conn(e,a)+i] * R^e_int[a,i](u; mat_p), original element order: dL/dtheta,
 * since every library energy is linear in its theta-scaled parameters (R^e_int(u; p(theta)) =
 * theta_e R^e_int(u; mat_p)).  Replaces `adjoint.py:75-106` (compute_implicit_vjp). */
int dfea_adjoint_grad(const dfea_desc* table, const int32_t* items, int n_items, const double* u,
                      const double* lam, double* g, int nen, void* stream);

#ifdef __cplusplus
}
#endif
#endif
