// libdfea_hip: gfx950 kernels and the C ABI (include/dfea.h) of the matrix-free Newton-Krylov hot path.
//
// Execution model
//   * Elements are grouped into chunks (<= CHUNK_LIN / CHUNK_QUAD elements, <= MAXLN distinct nodes).  One
//     workgroup handles one chunk: it gathers the chunk's nodal values into LDS, each thread evaluates one
//     element from structure-of-arrays moduli / geometry planes (coalesced HBM reads), and every node sums
//     the element vectors that touch it in a fixed order (elements are sorted by colour inside a chunk): the
//     product and PCG kernels through the chunk's inverse map (node -> element-vector slots, one barrier),
//     the other kernels colour by colour into LDS accumulators (elements of one colour share no node).  No
//     atomics, bit-reproducible results.
//   * Partials plans (the default): nodes used by several chunks receive one partial per chunk, stored
//     contiguously per node in chunk order, and the next kernel sums them in that order.  Halo plans
//     (quad4, optional): every node is owned by one chunk, and a chunk also evaluates the neighbouring
//     ("halo") elements that touch its owned nodes, so every owned node is completed inside its workgroup.
//   * PCG runs one kernel per iteration (Chronopoulos-Gear recurrences, see k_pcg), launched in pairs of
//     parities so that every buffer a launch writes is one no workgroup of that launch reads.
//   * Every launch processes a list of work items {solve, block}: a batch of independent solves of the
//     same element type and precision goes through one launch; per-solve reductions depend only on the
//     solve's own blocks, so a solve gives the same bits alone and inside a batch.
//   * Workgroup -> item mapping is XCD-aware: consecutive items (spatially adjacent chunks) land on the
//     same XCD, so nodal values shared by neighbouring chunks are reused in that XCD's L2.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstddef>
#include <mutex>
#include <type_traits>

#include "dfea.h"
#include "dfea_material.h"

namespace {

constexpr int CHUNK_LIN = 256;    // core elements per chunk for the linear elements (tri3, quad4)
constexpr int CHUNK_QUAD = 64;    // core elements per chunk for the quadratic elements (tri6, quad8, quad9)
constexpr int BS_LIN = 256;       // workgroup size for linear elements, partials plans
constexpr int BS_Q4H = 320;       // workgroup size for quad4 halo plans: core + halo elements (16x16 + 33)
constexpr int BS_QUAD = 64;       // workgroup size for quadratic elements (and small-chunk quad4 plans)
constexpr int BS_Q4S = 256;       // product and PCG workgroups of small-chunk quad4 plans: 4 threads per element
// workgroup size of the quadratic elements' product and PCG kernels: Gauss-point groups of 64 threads (2 in
// FP64, where registers limit the waves per SIMD, 3 in FP32; measured fastest for quad9 on 200x200 and on the
// 64x256 O-grid)
template <typename T>
constexpr int bs_groups() { return std::is_same<T, double>::value ? 128 : 192; }
// the PCG workgroup of quadratic-element meshes of at most PCG_EARLY_ITEMS chunks (DFEA_WIDE: one workgroup per CU
// or fewer): more Gauss-point groups per element, so more waves share each CU's work
template <typename T>
constexpr int bs_groups_wide() { return std::is_same<T, double>::value ? 192 : 256; }

// element types by node count: tri3 (3), quad4 (4), tri6 (6), quad8 (8), quad9 (9)
template <int NEN>
constexpr bool quadratic() { return NEN > 4; }
// Gauss points of each type's rule (`src/difffea/elements.py:38-57`): tri3 1 (centroid), quad4 2x2, tri6 6
// (degree 4), quad8 and quad9 3x3
template <int NEN>
constexpr int nq_of() { return NEN == 3 ? 1 : NEN == 4 ? 4 : NEN == 6 ? 6 : 9; }
// PCG launches of at most this many workgroups (one per CU or fewer: latency-bound) load the first node's
// state before the scalar step
constexpr int PCG_EARLY_ITEMS = 256;
// PCG launches whose element data (moduli and gradients) exceed this many bytes load it with nontemporal hints
constexpr double PCG_NT_BYTES = 240.0 * (1 << 20);

// product launches of at most this many workgroups request the Gauss points one step ahead (linear elements):
// that variant holds 3 workgroups of 256 threads per CU, the other 4
constexpr int APPLY_AHEAD_ITEMS = 3 * 256;
// local nodes per chunk (LDS capacity): the planner's chunks stay far below it (structured quad4 289, quad9 289-305,
// Delaunay quad4 335, quad9 333 measured); the LDS it saves lets more workgroups share a CU
constexpr int MAXLN = 448;
constexpr int VBS = 256;          // workgroup size of the vector kernels
constexpr int VPT = 4;            // dofs per thread in the vector kernels
constexpr int VBLK = VBS * VPT;   // dofs per vector block
constexpr int FBS = 256;          // workgroup size of the shared-node kernels
constexpr int NXCD = 8;

thread_local char g_err[256] = "";

int fail(int code, const char* what) {
    snprintf(g_err, sizeof g_err, "%s", what);
    return code;
}

int hip_status(const char* where) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        snprintf(g_err, sizeof g_err, "%s: %s", where, hipGetErrorString(e));
        return -1000 - static_cast<int>(e);
    }
    return 0;
}

template <typename T> struct V2 { T x, y; };

__device__ __forceinline__ int xcd_remap(int b, int n) {
    // hardware places workgroup b on XCD b % 8; give each XCD a contiguous range of items
    if (n < NXCD) return b;
    int x = b % NXCD, pos = b / NXCD;
    int base = n / NXCD, rem = n % NXCD;
    return x * base + (x < rem ? x : rem) + pos;
}

// deterministic workgroup sum; the result is valid on thread 0
template <int BS>
__device__ double block_sum(double v, double* s_red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) s_red[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        for (int i = 0; i < BS / 64; ++i) t += s_red[i];
    }
    __syncthreads();
    return t;
}

// v of the lane that DPP control CTRL selects (both 32-bit halves moved; every lane has a source)
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(u & 0xffffffffu), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(u >> 32), CTRL, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffu), lane);
    const int hi = __builtin_amdgcn_readlane((int)(u >> 32), lane);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// sum over the 64 lanes of a wave with the same bits in every lane: in each row of 16 lanes, the quad
// butterflies (DPP quad_perm lane ^ 1, lane ^ 2), then row_mirror (quad q with quad 3 - q) and
// row_half_mirror (quad 0 with 1, 2 with 3) -- each step adds a pair of lanes that swap partners, so both hold
// the same sum (addition commutes exactly) -- then the four row sums, read from lanes 0, 16, 32 and 48, are
// added in that fixed order.  Register moves only (no LDS round trips as in a shuffle).
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_f64<0xb1>(v);      // quad_perm [1, 0, 3, 2]
    v += dpp_f64<0x4e>(v);      // quad_perm [2, 3, 0, 1]
    v += dpp_f64<0x140>(v);     // row_mirror
    v += dpp_f64<0x141>(v);     // row_half_mirror
    return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

// three workgroup sums at once: wave sums (wave_sum), then the waves' sums in wave order.  ALL: every thread
// computes the totals (same order, same bits) after one barrier, and s_red must not be written again before a
// later barrier; otherwise the totals are valid on thread 0 and a second barrier frees s_red.  s_red holds
// 3 * BS / 64 doubles.
template <int BS, bool ALL = false>
__device__ void block_sum3(double& a, double& b, double& c, double* s_red) {
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_red[3 * w] = a;
        s_red[3 * w + 1] = b;
        s_red[3 * w + 2] = c;
    }
    __syncthreads();
    if (ALL || threadIdx.x == 0) {
        double ta = 0.0, tb = 0.0, tc = 0.0;
        for (int i = 0; i < BS / 64; ++i) {
            ta += s_red[3 * i];
            tb += s_red[3 * i + 1];
            tc += s_red[3 * i + 2];
        }
        a = ta;
        b = tb;
        c = tc;
    }
    if (!ALL) __syncthreads();
}

// sum of n partials in a fixed order, identical in every workgroup that calls it; valid on thread 0
template <int BS>
__device__ double sum_partials(const double* p, int n, double* s_red) {
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += BS) v += p[i];
    return block_sum<BS>(v, s_red);
}

// broadcast a thread-0 value to the workgroup
__device__ double bcast0(double v, double* s_slot) {
    if (threadIdx.x == 0) *s_slot = v;
    __syncthreads();
    double r = *s_slot;
    __syncthreads();
    return r;
}

// ------------------------------------------------------------------------------------------------------
// neo-Hookean closed forms (plane strain), see `src/difffea/materials.py:65-73`:
//   P = mu F + (lam lnJ - mu) G,  G = F^-T
//   d2psi/dF_ij dF_kl = mu d_ik d_jl + lam G_ij G_kl + (mu - lam lnJ) G_il G_kj
// ------------------------------------------------------------------------------------------------------
struct Kin {
    double F[4];     // F00 F01 F10 F11
    double G[4];     // F^-T
    double lnJ;
};

__device__ __forceinline__ Kin kinematics(const double F00, const double F01, const double F10, const double F11) {
    Kin k;
    k.F[0] = F00; k.F[1] = F01; k.F[2] = F10; k.F[3] = F11;
    const double J = F00 * F11 - F01 * F10;
    const double iJ = 1.0 / J;
    k.G[0] = F11 * iJ;  k.G[1] = -F10 * iJ;
    k.G[2] = -F01 * iJ; k.G[3] = F00 * iJ;
    k.lnJ = log(J);
    return k;
}

// 10 unique moduli in the order (00,01,02,03,11,12,13,22,23,33) of the flat index m = 2i+j
__device__ __forceinline__ void moduli10(const Kin& k, double mu, double lam, double w, double* c) {
    const double a = mu - lam * k.lnJ;
    int t = 0;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
#pragma unroll
        for (int n = m; n < 4; ++n) {
            const int i = m >> 1, j = m & 1, kk = n >> 1, l = n & 1;
            double v = lam * k.G[m] * k.G[n] + a * k.G[2 * i + l] * k.G[2 * kk + j];
            if (i == kk && j == l) v += mu;
            c[t++] = w * v;
        }
    }
}

// index of unique modulus (m, n) in the 10-entry list
__device__ __forceinline__ constexpr int uidx(int m, int n) {
    return m <= n ? (m * 4 - (m * (m - 1)) / 2 + (n - m)) : (n * 4 - (n * (n - 1)) / 2 + (m - n));
}

// value of a constrained dof of A(v) = m Jv + (1 - m) v (`operators.py:380,410`): 0 * Jv + v, which is v for
// finite Jv and NaN where an inverted element made the moduli NaN, as in the reference
template <typename T>
__device__ __forceinline__ T pivot(T jv, T v) { return (T)0 * jv + v; }

// R = m (R_int - lf f) (`operators.py:213-220`): constrained dofs are 0 * (R_int - lf f), NaN where an
// inverted element made R_int NaN, as in the reference
__device__ __forceinline__ double masked(bool fixed, double r) { return fixed ? 0.0 * r : r; }

// the clamped Jacobi entry of a free dof (`operators.py:555`): max(a, 1e-6) with NaN kept, as torch.clamp does
__device__ __forceinline__ double clamp_diag(double a) { return a < 1e-6 ? 1e-6 : a; }

// the scalar fields of a chunk record; the colour table is read through a pointer (a local copy of the
// record indexed by a loop variable would be placed in scratch memory)
struct ChunkView {
    int e0, ne, k0, nk, nown, pbase, ncol, h0, nh;
    const int32_t* col;
    const int32_t* hcol;
};

__device__ __forceinline__ ChunkView chunk_view(const dfea_chunk* c) {
    ChunkView v;
    v.e0 = c->e0; v.ne = c->ne; v.k0 = c->k0; v.nk = c->nk; v.nown = c->nown; v.pbase = c->pbase; v.ncol = c->ncol;
    v.h0 = c->h0; v.nh = c->nh;
    v.col = c->col;
    v.hcol = c->hcol;
    return v;
}

// element of thread t: core threads 0..ne-1 take the chunk's contiguous elements, halo threads ne..ne+nh-1
// the halo list.  e = plane column of the thread's data (n_el + halo entry for halo threads), el = the
// element id; returns false for idle threads.  l[] receives the element's local node indices.
template <int NEN>
__device__ __forceinline__ bool thread_element(const dfea_desc& D, const ChunkView& ch, int t, int& e, int* l,
                                               int* el = nullptr) {
    if (t < ch.ne) {
        e = ch.e0 + t;
        if (el) *el = e;
#pragma unroll
        for (int a = 0; a < NEN; ++a) l[a] = D.lelem[a * D.n_el + e];
        return true;
    }
    const int h = t - ch.ne;
    if (h < ch.nh) {
        const int hh = ch.h0 + h;
        e = D.n_el + hh;
        if (el) *el = D.hel[hh];
#pragma unroll
        for (int a = 0; a < NEN; ++a) l[a] = D.hlelem[a * D.n_halo + hh];
        return true;
    }
    e = 0;
    if (el) *el = 0;
    return false;
}

__device__ __forceinline__ bool in_colour(const ChunkView& ch, int t, int c) {
    if (t < ch.ne) return t >= ch.col[c] && t < ch.col[c + 1];
    const int h = t - ch.ne;
    return h >= ch.hcol[c] && h < ch.hcol[c + 1];
}

// reduction partials (dfea_desc.red): [0, 6 n_chunks) the PCG kernel's chunk partials (two parities x
// {r.u, w.u, r.r} per chunk), then their totals (two parities x 3, large solves: see k_pcg_reduce), then n_finblk
// unused, then 3 n_vblk for the vector kernels
__device__ __forceinline__ double* pcg_red(const dfea_desc& D, int par) { return D.red + 3 * par * D.n_chunks; }
__device__ __forceinline__ double* pcg_tot(const dfea_desc& D, int par) { return D.red + 6 * D.n_chunks + 3 * par; }
__device__ __forceinline__ double* vec_red(const dfea_desc& D) { return D.red + 6 * D.n_chunks + 6 + D.n_finblk; }

// solves of more chunks than this have their PCG chunk partials summed by k_pcg_reduce between iterations
// (one workgroup per solve), instead of by every workgroup of the next iteration.  Measured PCG iteration, FP64 /
// FP32, summed in every workgroup -> presummed: 500x500 Q9 (4,096 chunks) 186.2 / 114.6 -> 173.4 / 100.5 us,
// 800x800 Q4 (2,500) 121.1 / 55.6 -> 121.6 / 54.3 us, 566x566 Q4 (1,252) 45.5 / 29.1 -> 48.5 / 31.5 us (slower)
constexpr int PCG_PRESUM_CHUNKS = 2048;
constexpr int PCG_REDUCE_GRID = 64;      // workgroups of k_pcg_reduce (each scans its share of the items)

// work item of this workgroup; items == NULL: one solve (descriptor 0), block = workgroup index
struct Item { int s, b; };
__device__ __forceinline__ Item load_item(const int* items) {
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    Item it;
    if (items) {
        it.s = items[2 * b];
        it.b = items[2 * b + 1];
    } else {
        it.s = 0;
        it.b = b;
    }
    return it;
}


// ======================================================================================================
// geometry: dN_dX = dN_dxi J^-1, w = gw det J   (`src/difffea/operators.py:85-120`)
// items: {s, chunk}; thread per element of the chunk
// ======================================================================================================
template <int NEN, int NQ, int BS>
__global__ __launch_bounds__(BS) void k_geometry(const dfea_desc* __restrict__ tab, const int* __restrict__ items) {
    const Item it = load_item(items);
    const dfea_desc& D = tab[it.s];
    const ChunkView ch = chunk_view(D.chunks + it.b);
    const int t = threadIdx.x;
    int l[NEN], e, el;
    if (!thread_element<NEN>(D, ch, t, e, l, &el)) return;   // halo threads fill their copy columns
    const int ne = D.n_cols;
    double X[NEN][2];
#pragma unroll
    for (int a = 0; a < NEN; ++a) {
        const int n = D.conn[a * D.n_el + el];
        X[a][0] = D.nodes[2 * n];
        X[a][1] = D.nodes[2 * n + 1];
    }
    for (int q = 0; q < NQ; ++q) {
        double J00 = 0, J01 = 0, J10 = 0, J11 = 0;   // J[i][j] = dX_i / dxi_j
#pragma unroll
        for (int a = 0; a < NEN; ++a) {
            const double g0 = D.dNdxi[(q * NEN + a) * 2], g1 = D.dNdxi[(q * NEN + a) * 2 + 1];
            J00 += X[a][0] * g0; J01 += X[a][0] * g1;
            J10 += X[a][1] * g0; J11 += X[a][1] * g1;
        }
        const double det = J00 * J11 - J01 * J10;
        const double id = 1.0 / det;
        const double I00 = J11 * id, I01 = -J01 * id, I10 = -J10 * id, I11 = J00 * id;   // J^-1
#pragma unroll
        for (int a = 0; a < NEN; ++a) {
            const double g0 = D.dNdxi[(q * NEN + a) * 2], g1 = D.dNdxi[(q * NEN + a) * 2 + 1];
            const double d0 = g0 * I00 + g1 * I10;
            const double d1 = g0 * I01 + g1 * I11;
            const size_t p0 = (size_t)((q * NEN + a) * 2) * ne + e;
            D.dN64[p0] = d0;
            D.dN64[p0 + ne] = d1;
            if (D.dN32) {
                D.dN32[p0] = (float)d0;
                D.dN32[p0 + ne] = (float)d1;
            }
        }
        D.w[(size_t)q * ne + e] = D.gw[q] * det;
    }
}

// ======================================================================================================
// chunk kernels: gather -> element -> colour accumulate -> owned write / shared partial
// ======================================================================================================
template <typename T>
struct ChunkLDS {
    V2<T> v[MAXLN];      // gathered nodal values (unmasked)
    V2<T> acc[MAXLN];    // nodal accumulators
    unsigned char bits[MAXLN];
    double red[16];
};

template <typename T, int NEN>
__device__ __forceinline__ void colour_accumulate(const ChunkView& ch, int t, const int* l, const V2<T>* fe,
                                                  V2<T>* acc) {
    for (int c = 0; c < ch.ncol; ++c) {
        if (in_colour(ch, t, c)) {
#pragma unroll
            for (int a = 0; a < NEN; ++a) {
                V2<T> s = acc[l[a]];
                s.x += fe[a].x;
                s.y += fe[a].y;
                acc[l[a]] = s;
            }
        }
        __syncthreads();
    }
}

// ---- tangent product  y = m J (m v) + (1 - m) v  (`operators.py:385-412`) ----------------------------------
// Shared by the product kernel and the PCG kernel: the thread's element, the first PF Gauss points' moduli
// and gradients loaded before the gather barrier (so the HBM stream starts at once instead of after the
// gather's dependent round trips), and the element product accumulated colour by colour into LDS.

// Gauss-point groups: a workgroup of BS = G * CE threads over a partials-plan chunk of CE elements gives each
// element G threads, thread t taking element t mod CE and the Gauss points [g NQ / G, (g+1) NQ / G) of group
// g = t / CE; the groups' element vectors are added in group order.  (G = 1: one thread per element; halo
// plans, whose workgroups are not a multiple of the chunk, always use G = 1.)
// chunk elements of a kernel instantiation: CHUNK_LIN / CHUNK_QUAD by element type, CHUNK_QUAD for the quad4
// small-chunk plans (DFEA_SMALL)
template <int NEN>
constexpr int default_ce() { return quadratic<NEN>() ? CHUNK_QUAD : CHUNK_LIN; }

template <int NEN, int BS, int CEX = default_ce<NEN>()>
struct Groups {
    static constexpr int CE = CEX;
    static constexpr int G = BS % CE == 0 ? BS / CE : 1;
};

template <int NQ, int G>
struct GaussSplit {
    static constexpr int QG = (NQ + G - 1) / G;     // Gauss points per group, at most
};

// Thread t's element: partials plans cut the internal element order into chunks of exactly CE elements (the
// host checks it), so the element follows from the workgroup index alone and its loads need not wait for the
// chunk record; halo plans read the chunk record.
template <int NEN, int BS>
__device__ __forceinline__ bool chunk_element(const dfea_desc& D, const ChunkView& ch, int blk, int t, int& e, int* l) {
    constexpr int CE = Groups<NEN, BS>::CE;
    if (BS % CE == 0) {
        e = blk * CE + t % CE;
        if (e >= D.n_el) return false;
#pragma unroll
        for (int a = 0; a < NEN; ++a) l[a] = D.lelem[a * D.n_el + e];
        return true;
    }
    return thread_element<NEN>(D, ch, t, e, l);
}

// NT: nontemporal loads (the element data is streamed once per product; measured faster for the plain
// product, slower inside the PCG kernel, whose node updates share the cache)
// Loads through global-address-space pointers (global_load rather than flat_load: no LDS-aperture check, and
// the wait counts of loads and LDS operations stay apart).  gload: element i of an array through a 32-bit byte
// offset from a uniform base, the scalar-base + one-VGPR-offset form (the host keeps n_el and n_nodes below
// 2^28, so i * sizeof(T) fits).  gload2: node pair i of a V2 array as one 2-wide vector load.
template <typename T> using gptr = const __attribute__((address_space(1))) T*;
template <bool NT, typename T>
__device__ __forceinline__ T gload(const T* base, uint32_t i) {
    gptr<T> p = (gptr<T>)((gptr<char>)(gptr<T>)base + i * (uint32_t)sizeof(T));
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <typename T>
__device__ __forceinline__ V2<T> gload2(const V2<T>* base, uint32_t i) {
    typedef T vec2 __attribute__((ext_vector_type(2)));
    const vec2 w = *((gptr<vec2>)(gptr<V2<T>>)base + i);
    return V2<T>{w.x, w.y};
}
// entry k of the local-node tables (lnode, pslot, prange) of the chunk of work item b in a partials plan: the
// tables hold each chunk at the fixed stride lstride, so the index needs only the chunk index (no chunk record);
// k is clamped into the chunk's padded range (padding entries are valid, and dropped by the callers)
__device__ __forceinline__ uint32_t lentry(const dfea_desc& D, int b, int k) {
    const int ls = D.lstride;
    return (uint32_t)(b * ls + (k < ls ? k : ls - 1));
}

template <bool NT, typename T>
__device__ __forceinline__ T stream_load(const T* p) {
    if (NT) return __builtin_nontemporal_load(p);
    return *p;
}

template <int NEN, int NQ, typename T, bool NT = false>
__device__ __forceinline__ void load_gauss_point(const T* dN, const T* C, int ne, int e, int q, T (&d)[NEN][2],
                                                 T (&c)[10]) {
    // plane base pointers are uniform (scalar registers); the only per-thread term is the element column e
#pragma unroll
    for (int a = 0; a < NEN; ++a) {
        d[a][0] = gload<NT>(dN + (size_t)((q * NEN + a) * 2) * ne, (uint32_t)e);
        d[a][1] = gload<NT>(dN + (size_t)((q * NEN + a) * 2 + 1) * ne, (uint32_t)e);
    }
#pragma unroll
    for (int k = 0; k < 10; ++k) c[k] = gload<NT>(C + (size_t)(k * NQ + q) * ne, (uint32_t)e);
}

// the first PF Gauss points of the thread's group
template <int NEN, int NQ, int BS, int PF, typename T, bool NT = false, int CEX = default_ce<NEN>()>
__device__ __forceinline__ void prefetch_gauss_points(const T* dN, const T* C, int ne, int e, int t,
                                                      T (&d)[GaussSplit<NQ, Groups<NEN, BS, CEX>::G>::QG][NEN][2],
                                                      T (&c)[GaussSplit<NQ, Groups<NEN, BS, CEX>::G>::QG][10]) {
    constexpr int G = Groups<NEN, BS, CEX>::G, CE = Groups<NEN, BS, CEX>::CE, QG = GaussSplit<NQ, G>::QG;
    const int g = G > 1 ? t / CE : 0;
    const int q0 = g * NQ / G, q1 = (g + 1) * NQ / G;
#pragma unroll
    for (int j = 0; j < (PF < QG ? PF : QG); ++j)
        if (q0 + j < q1) load_gauss_point<NEN, NQ, T, NT>(dN, C, ne, e, q0 + j, d[j], c[j]);
}

// LDS of the product and PCG kernels on partials plans.  acc holds the chunk's element vectors (slot
// e_local * NEN + a) instead of node accumulators; slot / offs hold the chunk's inverse map (node k's element
// vectors are slots slot[offs[k] .. offs[k+1])), so every node sums its own in one pass after one barrier.
// SLOTS: element-vector slots of a chunk (chunk elements x nodes per element)
template <typename T, int SLOTS>
struct ProductLDS {
    static constexpr bool kInverse = true;
    V2<T> v[MAXLN];
    V2<T> acc[SLOTS];
    unsigned char bits[MAXLN];
    uint16_t slot[SLOTS];
    int offs[MAXLN + 1];
    double red[16];
};

// the inverse-map tables of a chunk, requested into registers early (the loads return while other work is in
// flight) and stored into LDS before the gather barrier; indices clamped into the chunk's ranges
template <int NEN, int BS, int CEX = default_ce<NEN>()>
struct InvRegs {
    static constexpr int CE = Groups<NEN, BS, CEX>::CE;
    static constexpr int NSL = (CE * NEN + BS - 1) / BS;     // slots per thread
    static constexpr int NOF = (MAXLN + 1 + BS - 1) / BS;    // node offsets per thread
    uint16_t sl[NSL];
    int of[NOF];
};

// FIXED: the offsets are addressed by the fixed stride of the local-node tables (no wait for the chunk record)
template <int NEN, int BS, int CEX = default_ce<NEN>(), bool FIXED = false>
__device__ __forceinline__ void inv_load(const dfea_desc& D, const ChunkView& ch, int blk, int t,
                                         InvRegs<NEN, BS, CEX>& R) {
    const int n = ch.ne * NEN;
    const uint32_t base = (uint32_t)(ch.e0 * NEN);
    const int ls = D.lstride;
#pragma unroll
    for (int i = 0; i < InvRegs<NEN, BS, CEX>::NSL; ++i) {
        const int j = t + i * BS;
        R.sl[i] = gload<false>(D.inv_slot, base + (uint32_t)(j < n ? j : n - 1));
    }
#pragma unroll
    for (int i = 0; i < InvRegs<NEN, BS, CEX>::NOF; ++i) {
        const int k = t + i * BS;
        // offsets k <= nk of the chunk (entries past nk hold the chunk's end)
        const uint32_t j = FIXED ? (uint32_t)(blk * ls + (k <= ls ? k : ls)) : (uint32_t)(ch.k0 + (k <= ch.nk ? k : ch.nk));
        R.of[i] = gload<false>(D.inv_off, j) - (int)base;
    }
}

template <int NEN, int BS, typename T, int CEX = default_ce<NEN>()>
__device__ __forceinline__ void inv_store(const ChunkView& ch, int t, const InvRegs<NEN, BS, CEX>& R,
                                          ProductLDS<T, CEX * NEN>& L) {
    const int n = ch.ne * NEN;
#pragma unroll
    for (int i = 0; i < InvRegs<NEN, BS, CEX>::NSL; ++i) {
        const int j = t + i * BS;
        if (j < n) L.slot[j] = R.sl[i];
    }
#pragma unroll
    for (int i = 0; i < InvRegs<NEN, BS, CEX>::NOF; ++i) {
        const int k = t + i * BS;
        if (k <= ch.nk) L.offs[k] = R.of[i];
    }
}

// the element vectors of local node k summed in slot (= colour) order: the same additions, in the same
// order, as the colour-by-colour accumulation into a zeroed accumulator
template <typename T, int SLOTS>
__device__ __forceinline__ V2<T> node_sum(const ProductLDS<T, SLOTS>& L, int k) {
    V2<T> s{0, 0};
    const int j1 = L.offs[k + 1];
    for (int j = L.offs[k]; j < j1; ++j) {
        const V2<T> f = L.acc[L.slot[j]];
        s.x += f.x;
        s.y += f.y;
    }
    return s;
}

// LDS for the element vectors of groups 1..G-1
template <int NEN, int BS, typename T, int CEX = default_ce<NEN>()>
struct GroupLDS {
    static constexpr int G = Groups<NEN, BS, CEX>::G, CE = Groups<NEN, BS, CEX>::CE;
    V2<T> fe[G > 1 ? G - 1 : 1][G > 1 ? CE : 1][NEN];
};

// fe = K_e (m v_e) from the gathered LDS values (the group's Gauss points, the groups added in order), then
// the colour-ordered sum into L.acc.  d / c hold the group's first PF Gauss points on entry.
template <int NEN, int NQ, int BS, int PF, typename T, bool NT = false, bool AHEAD = false,
          int CEX = default_ce<NEN>(), typename LDS>
__device__ __forceinline__ void element_product(const ChunkView& ch, int t, bool has_el, const int* l, const T* dN,
                                                const T* C, int ne, int e,
                                                T (&d)[GaussSplit<NQ, Groups<NEN, BS, CEX>::G>::QG][NEN][2],
                                                T (&c)[GaussSplit<NQ, Groups<NEN, BS, CEX>::G>::QG][10],
                                                LDS& L, GroupLDS<NEN, BS, T, CEX>& GL) {
    constexpr int G = Groups<NEN, BS, CEX>::G, CE = Groups<NEN, BS, CEX>::CE, QG = GaussSplit<NQ, G>::QG;
    const int g = G > 1 ? t / CE : 0;
    const int q0 = g * NQ / G, q1 = (g + 1) * NQ / G;
    V2<T> fe[NEN];
#pragma unroll
    for (int a = 0; a < NEN; ++a) fe[a] = V2<T>{0, 0};
    if (has_el) {
        V2<T> ve[NEN];
#pragma unroll
        for (int a = 0; a < NEN; ++a) {
            ve[a] = L.v[l[a]];
            const unsigned char b = L.bits[l[a]];
            if (b & 1) ve[a].x = 0;
            if (b & 2) ve[a].y = 0;
        }
#pragma unroll
        for (int j = 0; j < QG; ++j) {
            const int q = q0 + j;
            if (q >= q1) break;
            // AHEAD: the Gauss point PF ahead is requested before this one is used, so PF points stay in flight
            // (costs registers; used where they are plentiful)
            if (AHEAD) {
                if (j + PF < QG && q + PF < q1) load_gauss_point<NEN, NQ, T, NT>(dN, C, ne, e, q + PF, d[j + PF], c[j + PF]);
            } else if (j >= PF) {
                load_gauss_point<NEN, NQ, T, NT>(dN, C, ne, e, q, d[j], c[j]);
            }
            T gr[4] = {0, 0, 0, 0};    // gr[2i+k] = sum_a ve[a][i] dN[a][k]
#pragma unroll
            for (int a = 0; a < NEN; ++a) {
                gr[0] += ve[a].x * d[j][a][0];
                gr[1] += ve[a].x * d[j][a][1];
                gr[2] += ve[a].y * d[j][a][0];
                gr[3] += ve[a].y * d[j][a][1];
            }
            T sq[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                T acc = 0;
#pragma unroll
                for (int n = 0; n < 4; ++n) acc += c[j][uidx(m, n)] * gr[n];
                sq[m] = acc;
            }
#pragma unroll
            for (int a = 0; a < NEN; ++a) {
                fe[a].x += sq[0] * d[j][a][0] + sq[1] * d[j][a][1];
                fe[a].y += sq[2] * d[j][a][0] + sq[3] * d[j][a][1];
            }
        }
    }
    if (G > 1) {
        if (g > 0 && has_el) {
#pragma unroll
            for (int a = 0; a < NEN; ++a) GL.fe[g - 1][t % CE][a] = fe[a];
        }
        __syncthreads();
        if (g == 0 && has_el) {
            for (int h = 1; h < G; ++h) {
#pragma unroll
                for (int a = 0; a < NEN; ++a) {
                    const V2<T> o = GL.fe[h - 1][t][a];
                    fe[a].x += o.x;
                    fe[a].y += o.y;
                }
            }
        }
    }
    if constexpr (std::is_same<LDS, ProductLDS<T, CE * NEN>>::value) {
        // inverse-map sums: element vectors into their slots, one barrier, then every node sums its own
        if (g == 0 && has_el) {
#pragma unroll
            for (int a = 0; a < NEN; ++a) L.acc[(t % CE) * NEN + a] = fe[a];
        }
        __syncthreads();
    } else {
        colour_accumulate<T, NEN>(ch, t, l, fe, L.acc);
    }
}

// PF: Gauss points whose data is loaded before the gather barrier; MINW: minimum waves per SIMD the
// register allocation must allow (1 = no constraint); STREAM_FIRST (partials plans): those PF Gauss points are
// requested before the gather rather than behind it; DIRECT (one solve): the chunk records come as the first
// argument (preloaded in registers), so the record load does not wait for the descriptor load
template <int NEN, int NQ, int BS, typename T, int PF = 2, int MINW = 1, bool AHEAD = false,
          bool STREAM_FIRST = false, int CEX = default_ce<NEN>(), bool DIRECT = false>
__global__ __launch_bounds__(BS, MINW) void k_apply(const dfea_chunk* __restrict__ chunks_arg,
                                                   const dfea_desc* __restrict__ tab, const int* __restrict__ items,
                                                   const T* __restrict__ vin, T* __restrict__ vout,
                                                   V2<T>* __restrict__ part_arg, int part_cap) {
    // partials plans (chunks of exactly CE elements) sum nodes through the inverse map, halo plans by colour
    constexpr bool INV = BS % Groups<NEN, BS, CEX>::CE == 0;
    __shared__ std::conditional_t<INV, ProductLDS<T, CEX * NEN>, ChunkLDS<T>> L;
    __shared__ GroupLDS<NEN, BS, T, CEX> GL;
    const Item it = load_item(items);
    const dfea_desc& D = tab[it.s];
    const ChunkView ch = chunk_view((DIRECT ? chunks_arg : D.chunks) + it.b);
    const int t = threadIdx.x;
    const T* v = vin ? vin : static_cast<const T*>(D.vec[DFEA_P]);
    T* y = vout ? vout : static_cast<T*>(D.vec[DFEA_W]);
    const int ne = D.n_cols;
    const T* dN = std::is_same<T, double>::value ? (const T*)D.dN64 : (const T*)D.dN32;
    const T* C = std::is_same<T, double>::value ? (const T*)D.C64 : (const T*)D.C32;

    // 0. element (local node indices).  Partials plans: threads past the chunk's elements read its last element
    // and drop the result, so every load below is unconditional (no branch splits the loads in flight).  The
    // element range comes from the chunk record (tail-balanced plans have chunks of two sizes), except where the
    // element stream starts before the record arrives (STREAM_FIRST: plans of equal chunks only)
    constexpr int CE = Groups<NEN, BS, CEX>::CE;
    int l[NEN];
    int e;
    bool has_el;
    if constexpr (BS % CE == 0 && STREAM_FIRST) {
        e = it.b * CE + t % CE;
        has_el = e < D.n_el;
        if (!has_el) e = D.n_el - 1;
    } else if constexpr (BS % CE == 0) {
        const int tc = t % CE;
        has_el = tc < ch.ne;
        e = ch.e0 + (has_el ? tc : ch.ne - 1);
    } else {
        has_el = thread_element<NEN>(D, ch, t, e, l);
    }
    constexpr int QG = GaussSplit<NQ, Groups<NEN, BS, CEX>::G>::QG;
    T d[QG][NEN][2];
    T c[QG][10];
    if constexpr (INV && STREAM_FIRST) {
        // the bulk stream starts right after the descriptor; the gather below then returns behind it
        prefetch_gauss_points<NEN, NQ, BS, PF, T, true, CEX>(dN, C, ne, e, t, d, c);
        __builtin_amdgcn_sched_barrier(0);
    }

    // 1. gather: node ids, then the nodal values; the element data of the first PF Gauss points is requested
    // behind them (unless STREAM_FIRST), so the gather's loads return first
    constexpr int NG = (MAXLN + BS - 1) / BS;
    uint32_t gid[NG];
    V2<T> gv[NG];
#pragma unroll
    for (int i = 0; i < NG; ++i) {
        const int k = t + i * BS;
        gid[i] = gload<false>(D.lnode, (uint32_t)(ch.k0 + (k < ch.nk ? k : ch.nk - 1)));
    }
#pragma unroll
    for (int i = 0; i < NG; ++i) gv[i] = gload2<T>(reinterpret_cast<const V2<T>*>(v), gid[i] & DFEA_NODE_MASK);
    // partial slot of every shared entry the thread writes at the end (loaded now, used after the product)
    int slot[NG];
#pragma unroll
    for (int i = 0; i < NG; ++i) {
        const int k = t + i * BS;
        const int ent = k >= ch.nown && k < ch.nk ? ch.pbase + (k - ch.nown) : 0;
        slot[i] = D.n_part ? gload<false>(D.pslot, (uint32_t)ent) : 0;
    }
    InvRegs<NEN, BS, CEX> inv;
    if constexpr (INV) {
#pragma unroll
        for (int a = 0; a < NEN; ++a) l[a] = gload<false>(D.lelem, (uint32_t)(a * D.n_el + e));
        inv_load<NEN, BS, CEX>(D, ch, it.b, t, inv);
    }
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (!(INV && STREAM_FIRST)) prefetch_gauss_points<NEN, NQ, BS, PF, T, true, CEX>(dN, C, ne, e, t, d, c);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < NG; ++i) {
        const int k = t + i * BS;
        if (k < ch.nk) {
            L.v[k] = gv[i];
            if constexpr (!INV) L.acc[k] = V2<T>{0, 0};
            L.bits[k] = (unsigned char)(gid[i] >> 30);
        }
    }
    if constexpr (INV) inv_store<NEN, BS, T, CEX>(ch, t, inv, L);
    __builtin_amdgcn_sched_barrier(0);
    __syncthreads();
    __builtin_amdgcn_sched_barrier(0);

    // 2. element products; node sums in colour order (inverse map or colour passes)
    element_product<NEN, NQ, BS, PF, T, true, AHEAD, CEX>(ch, t, has_el, l, dN, C, ne, e, d, c, L, GL);

    // 3. owned nodes: final value; shared nodes (partials plans): partial (into part_arg when given: the
    // caller's buffer of part_cap entries, every slot checked against it)
    V2<T>* part = part_arg ? part_arg : reinterpret_cast<V2<T>*>(D.part);
#pragma unroll
    for (int i = 0; i < NG; ++i) {
        const int k = t + i * BS;
        if (k >= ch.nk) break;
        V2<T> acc;
        if constexpr (INV) acc = node_sum<T>(L, k);
        else acc = L.acc[k];
        if (k < ch.nown) {
            const V2<T> raw = L.v[k];
            V2<T> out;
            out.x = (gid[i] & DFEA_FIXED_X) ? pivot(acc.x, raw.x) : acc.x;
            out.y = (gid[i] & DFEA_FIXED_Y) ? pivot(acc.y, raw.y) : acc.y;
            reinterpret_cast<V2<T>*>(y)[(int)(gid[i] & DFEA_NODE_MASK)] = out;
        } else if (D.n_part && slot[i] < part_cap) {
            part[slot[i]] = acc;
        }
    }
}

// sum of the cnt partials of a shared node starting at slot j0, in slot order; the first SUM_UNROLL are
// loaded together, so the loads are in flight at once rather than one after another
constexpr int SUM_UNROLL = 4;
template <typename T>
__device__ __forceinline__ V2<T> sum_slots(const V2<T>* part, int j0, int cnt) {
    V2<T> pv[SUM_UNROLL];
#pragma unroll
    for (int q = 0; q < SUM_UNROLL; ++q)
        if (q < cnt) pv[q] = part[j0 + q];
    V2<T> acc{0, 0};
#pragma unroll
    for (int q = 0; q < SUM_UNROLL; ++q) {
        if (q < cnt) {
            acc.x += pv[q].x;
            acc.y += pv[q].y;
        }
    }
    for (int j = j0 + SUM_UNROLL; j < j0 + cnt; ++j) {
        const V2<T> pj = part[j];
        acc.x += pj.x;
        acc.y += pj.y;
    }
    return acc;
}

// shared nodes of the tangent product (partials plans): partials summed in slot order
template <typename T>
__global__ __launch_bounds__(FBS) void k_apply_fin(const dfea_desc* __restrict__ tab, const int* __restrict__ items,
                                                  const T* __restrict__ vin, T* __restrict__ vout,
                                                  const V2<T>* __restrict__ part_arg, int part_cap) {
    const Item it = load_item(items);
    const dfea_desc& D = tab[it.s];
    const int i = it.b * FBS + threadIdx.x;
    const T* v = vin ? vin : static_cast<const T*>(D.vec[DFEA_P]);
    T* y = vout ? vout : static_cast<T*>(D.vec[DFEA_W]);
    if (i < D.n_shared) {
        const V2<T>* part = part_arg ? part_arg : reinterpret_cast<const V2<T>*>(D.part);
        const int j0 = D.sh_off[i];
        const int j1 = D.sh_off[i + 1];
        if (j1 > part_cap) return;
        const V2<T> acc = sum_slots<T>(part, j0, j1 - j0);
        const uint32_t g = D.sh_node[i];
        const int n = (int)(g & DFEA_NODE_MASK);
        const V2<T> raw = reinterpret_cast<const V2<T>*>(v)[n];
        V2<T> out;
        out.x = (g & DFEA_FIXED_X) ? pivot(acc.x, raw.x) : acc.x;
        out.y = (g & DFEA_FIXED_Y) ? pivot(acc.y, raw.y) : acc.y;
        reinterpret_cast<V2<T>*>(y)[n] = out;
    }
}

// the same for one solve, with every pointer an argument (kernel arguments arrive preloaded in SGPRs): the node
// tables' loads start at once instead of after the descriptor load; part holds at least sh_off[n_shared] entries
template <typename T>
__global__ __launch_bounds__(FBS) void k_apply_fin_direct(const int* __restrict__ sh_off,
                                                         const uint32_t* __restrict__ sh_node,
                                                         const V2<T>* __restrict__ part, const T* __restrict__ v,
                                                         T* __restrict__ y, int n_shared) {
    const int i = xcd_remap(blockIdx.x, gridDim.x) * FBS + threadIdx.x;
    if (i < n_shared) {
        const int j0 = sh_off[i];
        const int j1 = sh_off[i + 1];
        const V2<T> acc = sum_slots<T>(part, j0, j1 - j0);
        const uint32_t g = sh_node[i];
        const int n = (int)(g & DFEA_NODE_MASK);
        const V2<T> raw = reinterpret_cast<const V2<T>*>(v)[n];
        V2<T> out;
        out.x = (g & DFEA_FIXED_X) ? pivot(acc.x, raw.x) : acc.x;
        out.y = (g & DFEA_FIXED_Y) ? pivot(acc.y, raw.y) : acc.y;
        reinterpret_cast<V2<T>*>(y)[n] = out;
    }
}

// w P(F) of the element's material (the neo-Hookean closed form inline, the library through dfea_mat)
template <bool LIB>
__device__ __forceinline__ void piola_w(const dfea_mat::Mat& M, double F00, double F01, double F10, double F11,
                                        double w, double* P) {
    if (!LIB || M.id == DFEA_MAT_NEO_HOOKEAN) {
        const Kin k = kinematics(F00, F01, F10, F11);
        const double mu = M.p[0], b = M.p[1] * k.lnJ - mu;
#pragma unroll
        for (int m = 0; m < 4; ++m) P[m] = w * (mu * k.F[m] + b * k.G[m]);
        return;
    }
    const double F[4] = {F00, F01, F10, F11};
    dfea_mat::piola<LIB>(M, F, P);
#pragma unroll
    for (int m = 0; m < 4; ++m) P[m] *= w;
}

// ---- residual  R = m (R_int(u) - lf f)  (`operators.py:131-220`) -------------------------------------------
// REACT: support reactions instead (`postprocess.py:121-137`): the chunk's sums of the internal force over its
// fixed dofs (owned nodes, and its partial of each shared node: the total is linear in the partials) go to
// red[b] (x) and red[n_chunks + b] (y); R is not written
template <int NEN, int NQ, int BS, bool LIB, bool REACT = false>
__global__ __launch_bounds__(BS) void k_residual(const dfea_desc* __restrict__ tab, const int* __restrict__ items,
                                                const double* __restrict__ uin, double* __restrict__ rout,
                                                double lf_arg) {
    __shared__ ChunkLDS<double> L;
    const Item it = load_item(items);
    const dfea_desc& D = tab[it.s];
    const ChunkView ch = chunk_view(D.chunks + it.b);
    const int t = threadIdx.x;
    const double* u = uin ? uin : static_cast<const double*>(D.vec[DFEA_U]);
    double* R = rout ? rout : static_cast<double*>(D.vec[DFEA_RES]);
    const double lf = lf_arg >= 0.0 ? lf_arg : D.sc[DFEA_SC_LF];
    for (int k = t; k < ch.nk; k += BS) {
        const uint32_t g = D.lnode[ch.k0 + k];
        const int n = (int)(g & DFEA_NODE_MASK);
        L.v[k] = reinterpret_cast<const V2<double>*>(u)[n];
        L.acc[k] = V2<double>{0, 0};
        L.bits[k] = (unsigned char)(g >> 30);
    }
    __syncthreads();
    int l[NEN];
    int e;
    const bool has_el = thread_element<NEN>(D, ch, t, e, l);
    V2<double> fe[NEN];
    if (has_el) {
        const int ne = D.n_cols;
        const dfea_mat::Mat M = dfea_mat::element_material<LIB>(D, D.theta[e]);
        V2<double> ue[NEN];
#pragma unroll
        for (int a = 0; a < NEN; ++a) {
            ue[a] = L.v[l[a]];
            fe[a] = V2<double>{0, 0};
        }
        for (int q = 0; q < NQ; ++q) {
            double d[NEN][2];
            double F00 = 1, F01 = 0, F10 = 0, F11 = 1;
#pragma unroll
            for (int a = 0; a < NEN; ++a) {
                d[a][0] = D.dN64[(size_t)((q * NEN + a) * 2) * ne + e];
                d[a][1] = D.dN64[(size_t)((q * NEN + a) * 2 + 1) * ne + e];
                F00 += ue[a].x * d[a][0]; F01 += ue[a].x * d[a][1];
                F10 += ue[a].y * d[a][0]; F11 += ue[a].y * d[a][1];
            }
            const double w = D.w[(size_t)q * ne + e];
            double P[4];
            piola_w<LIB>(M, F00, F01, F10, F11, w, P);
#pragma unroll
            for (int a = 0; a < NEN; ++a) {
                fe[a].x += P[0] * d[a][0] + P[1] * d[a][1];
                fe[a].y += P[2] * d[a][0] + P[3] * d[a][1];
            }
        }
    }
    colour_accumulate<double, NEN>(ch, t, l, fe, L.acc);
    if constexpr (REACT) {
        double rx = 0.0, ry = 0.0, unused = 0.0;
        for (int k = t; k < ch.nk; k += BS) {
            if (k >= ch.nown && !D.n_part) continue;     // halo plans: another chunk owns the node
            const V2<double> acc = L.acc[k];
            const unsigned char b = L.bits[k];
            if (b & 1) rx += acc.x;
            if (b & 2) ry += acc.y;
        }
        block_sum3<BS>(rx, ry, unused, L.red);
        if (t == 0) {
            D.red[it.b] = rx;
            D.red[D.n_chunks + it.b] = ry;
        }
        return;
    }
    V2<double>* part = reinterpret_cast<V2<double>*>(D.part);
    for (int k = t; k < ch.nk; k += BS) {
        const V2<double> acc = L.acc[k];
        if (k < ch.nown) {
            const int n = (int)(D.lnode[ch.k0 + k] & DFEA_NODE_MASK);
            const unsigned char b = L.bits[k];
            V2<double> out;
            out.x = masked(b & 1, acc.x - lf * D.f[2 * n]);
            out.y = masked(b & 2, acc.y - lf * D.f[2 * n + 1]);
            reinterpret_cast<V2<double>*>(R)[n] = out;
        } else if (D.n_part) {
            part[D.pslot[ch.pbase + (k - ch.nown)]] = acc;
        }
    }
}

__global__ __launch_bounds__(FBS) void k_residual_fin(const dfea_desc* __restrict__ tab, const int* __restrict__ items,
                                                     double* __restrict__ rout, double lf_arg) {
    const Item it = load_item(items);
    const dfea_desc& D = tab[it.s];
    const int i = it.b * FBS + threadIdx.x;
    if (i >= D.n_shared) return;
    double* R = rout ? rout : static_cast<double*>(D.vec[DFEA_RES]);
    const double lf = lf_arg >= 0.0 ? lf_arg : D.sc[DFEA_SC_LF];
    const V2<double>* part = reinterpret_cast<const V2<double>*>(D.part);
    const int j0 = D.sh_off[i];
    const V2<double> acc = sum_slots<double>(part, j0, D.sh_off[i + 1] - j0);
    const uint32_t g = D.sh_node[i];
    const int n = (int)(g & DFEA_NODE_MASK);
    V2<double> out;
    out.x = masked(g & DFEA_FIXED_X, acc.x - lf * D.f[2 * n]);
    out.y = masked(g & DFEA_FIXED_Y, acc.y - lf * D.f[2 * n + 1]);
    reinterpret_cast<V2<double>*>(R)[n] = out;
}

// ---- moduli + preconditioner  (`operators.py:316-342`, `:477-556`, `:588-638`) -------------------------------
// the preconditioner of one node from its summed diagonal entries a = (A_xx, A_yy) and, for block Jacobi, the
// off-diagonal entry axy of its 2x2 block: point Jacobi keeps the clamped diagonal (1 on fixed dofs); block
// Jacobi pivots fixed dofs (identity row and column), adds 1e-12 I and stores the closed-form inverse,
// {m00, m11} at the node's dofs and m01 after the 2N diagonal entries
template <typename CT>
__device__ __forceinline__ void store_precond(const dfea_desc& D, int n, unsigned char b, V2<double> a, double axy) {
    CT* out = std::is_same<CT, double>::value ? (CT*)D.diag64 : (CT*)D.diag32;
    double d0, d1;
    if (D.precond == DFEA_PRECOND_BLOCK) {
        const double b00 = ((b & 1) ? 1.0 : a.x) + 1e-12;
        const double b11 = ((b & 2) ? 1.0 : a.y) + 1e-12;
        const double b01 = b ? 0.0 : axy;
        const double det = b00 * b11 - b01 * b01;
        d0 = b11 / det;
        d1 = b00 / det;
        out[2 * D.n_nodes + n] = (CT)(-b01 / det);
    } else {
        d0 = (b & 1) ? 1.0 : clamp_diag(a.x);
        d1 = (b & 2) ? 1.0 : clamp_diag(a.y);
    }
    out[2 * n] = (CT)d0;
    out[2 * n + 1] = (CT)d1;
}

template <int NEN, int NQ, int BS, typename CT, bool LIB>
__global__ __launch_bounds__(BS) void k_moduli(const dfea_desc* __restrict__ tab, const int* __restrict__ items,
                                              const double* __restrict__ uin) {
    __shared__ ChunkLDS<double> L;
    const Item it = load_item(items);
    const dfea_desc& D = tab[it.s];
    const ChunkView ch = chunk_view(D.chunks + it.b);
    const int t = threadIdx.x;
    const double* u = uin ? uin : static_cast<const double*>(D.vec[DFEA_U]);
    for (int k = t; k < ch.nk; k += BS) {
        const uint32_t g = D.lnode[ch.k0 + k];
        const int n = (int)(g & DFEA_NODE_MASK);
        L.v[k] = reinterpret_cast<const V2<double>*>(u)[n];
        L.acc[k] = V2<double>{0, 0};
        L.bits[k] = (unsigned char)(g >> 30);
    }
    __syncthreads();
    int l[NEN];
    int e;
    const bool has_el = thread_element<NEN>(D, ch, t, e, l);
    const bool blk = D.precond == DFEA_PRECOND_BLOCK;
    V2<double> fe[NEN];
    V2<double> fo[NEN];     // block Jacobi: {A_xy, 0} of the element's nodal blocks
    if (has_el) {
        const int ne = D.n_cols;
        const dfea_mat::Mat M = dfea_mat::element_material<LIB>(D, D.theta[e]);
        CT* Cout = std::is_same<CT, double>::value ? (CT*)D.C64 : (CT*)D.C32;
        V2<double> ue[NEN];
#pragma unroll
        for (int a = 0; a < NEN; ++a) {
            ue[a] = L.v[l[a]];
            fe[a] = V2<double>{0, 0};
            fo[a] = V2<double>{0, 0};
        }
        for (int q = 0; q < NQ; ++q) {
            double d[NEN][2];
            double F00 = 1, F01 = 0, F10 = 0, F11 = 1;
#pragma unroll
            for (int a = 0; a < NEN; ++a) {
                d[a][0] = D.dN64[(size_t)((q * NEN + a) * 2) * ne + e];
                d[a][1] = D.dN64[(size_t)((q * NEN + a) * 2 + 1) * ne + e];
                F00 += ue[a].x * d[a][0]; F01 += ue[a].x * d[a][1];
                F10 += ue[a].y * d[a][0]; F11 += ue[a].y * d[a][1];
            }
            double c[10];
            const double w = D.w[(size_t)q * ne + e];
            if (!LIB || M.id == DFEA_MAT_NEO_HOOKEAN) {
                moduli10(kinematics(F00, F01, F10, F11), M.p[0], M.p[1], w, c);
            } else {
                const double F[4] = {F00, F01, F10, F11};
                dfea_mat::moduli<LIB>(M, F, w, c);
            }
#pragma unroll
            for (int kk = 0; kk < 10; ++kk) Cout[(size_t)(kk * NQ + q) * ne + e] = (CT)c[kk];   // halo: copy column
            // K^e_{(a,i),(a,i)} += sum_{j,l} C[2i+j][2i+l] dN[a][j] dN[a][l]
#pragma unroll
            for (int a = 0; a < NEN; ++a) {
                const double d0 = d[a][0], d1 = d[a][1];
                fe[a].x += c[uidx(0, 0)] * d0 * d0 + 2.0 * c[uidx(0, 1)] * d0 * d1 + c[uidx(1, 1)] * d1 * d1;
                fe[a].y += c[uidx(2, 2)] * d0 * d0 + 2.0 * c[uidx(2, 3)] * d0 * d1 + c[uidx(3, 3)] * d1 * d1;
                // K^e_{(a,0),(a,1)} = sum_{j,l} C[j][2+l] dN[a][j] dN[a][l]
                if (blk)
                    fo[a].x += c[uidx(0, 2)] * d0 * d0 + (c[uidx(0, 3)] + c[uidx(1, 2)]) * d0 * d1 +
                               c[uidx(1, 3)] * d1 * d1;
            }
        }
    }
    colour_accumulate<double, NEN>(ch, t, l, fe, L.acc);
    if (blk) {      // the off-diagonal entries, accumulated in L.v once every thread has read its gathered values
        __syncthreads();
        for (int k = t; k < ch.nk; k += BS) L.v[k] = V2<double>{0, 0};
        __syncthreads();
        colour_accumulate<double, NEN>(ch, t, l, fo, L.v);
    }
    V2<double>* part = reinterpret_cast<V2<double>*>(D.part);
    for (int k = t; k < ch.nk; k += BS) {
        const V2<double> acc = L.acc[k];
        const double axy = blk ? L.v[k].x : 0.0;
        if (k < ch.nown) {
            const int n = (int)(D.lnode[ch.k0 + k] & DFEA_NODE_MASK);
            store_precond<CT>(D, n, L.bits[k], acc, axy);
        } else if (D.n_part) {
            const int slot = D.pslot[ch.pbase + (k - ch.nown)];
            part[slot] = acc;
            if (blk) part[D.n_part + slot] = V2<double>{axy, 0.0};
        }
    }
}

template <typename CT>
__global__ __launch_bounds__(FBS) void k_moduli_fin(const dfea_desc* __restrict__ tab, const int* __restrict__ items) {
    const Item it = load_item(items);
    const dfea_desc& D = tab[it.s];
    const int i = it.b * FBS + threadIdx.x;
    if (i >= D.n_shared) return;
    const V2<double>* part = reinterpret_cast<const V2<double>*>(D.part);
    const int j0 = D.sh_off[i];
    const V2<double> acc = sum_slots<double>(part, j0, D.sh_off[i + 1] - j0);
    const double axy = D.precond == DFEA_PRECOND_BLOCK ? sum_slots<double>(part + D.n_part, j0, D.sh_off[i + 1] - j0).x
                                                       : 0.0;
    const uint32_t g = D.sh_node[i];
    store_precond<CT>(D, (int)(g & DFEA_NODE_MASK), (unsigned char)(g >> 30), acc, axy);
}

// the sums of the reaction kernel's chunk partials, per solve: out[2 s], out[2 s + 1]
__global__ __launch_bounds__(VBS) void k_reactions_final(const dfea_desc* __restrict__ tab, const int* __restrict__ items,
                                                        double* __restrict__ out) {
    __shared__ double s_red[16];
    const Item it = load_item(items);
    const dfea_desc& D = tab[it.s];
    const double rx = sum_partials<VBS>(D.red, D.n_chunks, s_red);
    const double ry = sum_partials<VBS>(D.red + D.n_chunks, D.n_chunks, s_red);
    if (threadIdx.x == 0) {
        out[2 * it.s] = rx;
        out[2 * it.s + 1] = ry;
    }
}

// ---- post-processing fields  (`postprocess.py:24-118`) ---------------------------------------------------------
// Per element: the area sum_q w_q det J_q and the means over it (weights w_q det J_q) of J, of the Cauchy
// stress sigma = P3 F3^T / J of the plane-strain embedding F3 = diag(F, 1) -- for the neo-Hookean energy
// sigma = (mu (F F^T - I) + lam ln J I) / J in-plane and sigma_zz = lam ln J / J -- and of the energy density
// psi = mu/2 (F:F - 2 - 2 ln J) + lam/2 (ln J)^2; the von Mises stress of the mean stress.  Written to
// eout[k][original element id], k = area, J, sxx, syy, sxy, mises, psi.  Per node: the element values
// J, sxx, syy, sxy, mises, psi summed over the elements at the node and divided by their number, written to
// nout[k][node]; shared nodes leave four partial pairs per slot in work ([4][n_part] pairs) for k_fields_fin.
template <int NEN>
__device__ __forceinline__ void colour_accumulate4(const ChunkView& ch, int t, const int* l, const V2<double>* val,
                                                   V2<double>* a0, V2<double>* a1, V2<double>* a2, V2<double>* a3) {
    for (int c = 0; c < ch.ncol; ++c) {
        if (in_colour(ch, t, c)) {
#pragma unroll
            for (int a = 0; a < NEN; ++a) {
                const int k = l[a];
                a0[k].x += val[0].x; a0[k].y += val[0].y;
                a1[k].x += val[1].x; a1[k].y += val[1].y;
                a2[k].x += val[2].x; a2[k].y += val[2].y;
                a3[k].x += val[3].x; a3[k].y += val[3].y;
            }
        }
        __syncthreads();
    }
}

template <int NEN, int NQ, int BS, bool LIB>
__global__ __launch_bounds__(BS) void k_fields(const dfea_desc* __restrict__ tab, const int* __restrict__ items,
                                              const double* __restrict__ uin, double* __restrict__ eout,
                                              double* __restrict__ nout, double* __restrict__ work) {
    __shared__ ChunkLDS<double> L;
    __shared__ V2<double> acc1[MAXLN], acc2[MAXLN], acc3[MAXLN];
    const Item it = load_item(items);
    const dfea_desc& D = tab[it.s];
    const ChunkView ch = chunk_view(D.chunks + it.b);
    const int t = threadIdx.x;
    const double* u = uin ? uin : static_cast<const double*>(D.vec[DFEA_U]);
    for (int k = t; k < ch.nk; k += BS) {
        const int n = (int)(D.lnode[ch.k0 + k] & DFEA_NODE_MASK);
        L.v[k] = reinterpret_cast<const V2<double>*>(u)[n];
        L.acc[k] = acc1[k] = acc2[k] = acc3[k] = V2<double>{0, 0};
    }
    __syncthreads();
    int l[NEN];
    int e;
    const bool has_el = thread_element<NEN>(D, ch, t, e, l);
    V2<double> val[4] = {};
    if (has_el) {
        const int ne = D.n_cols;
        const dfea_mat::Mat M = dfea_mat::element_material<LIB>(D, D.theta[e]);
        V2<double> ue[NEN];
#pragma unroll
        for (int a = 0; a < NEN; ++a) ue[a] = L.v[l[a]];
        double area = 0, mJ = 0, mxx = 0, myy = 0, mxy = 0, mzz = 0, mpsi = 0;
        for (int q = 0; q < NQ; ++q) {
            double F00 = 1, F01 = 0, F10 = 0, F11 = 1;
#pragma unroll
            for (int a = 0; a < NEN; ++a) {
                const double d0 = D.dN64[(size_t)((q * NEN + a) * 2) * ne + e];
                const double d1 = D.dN64[(size_t)((q * NEN + a) * 2 + 1) * ne + e];
                F00 += ue[a].x * d0; F01 += ue[a].x * d1;
                F10 += ue[a].y * d0; F11 += ue[a].y * d1;
            }
            const double J = F00 * F11 - F01 * F10;
            const double w = D.w[(size_t)q * ne + e];
            const double F[4] = {F00, F01, F10, F11};
            double P[4], P33, psi;
            dfea_mat::stress_state<LIB>(M, F, P, P33, psi);
            // sigma = P3 F3^T / J
            area += w;
            mJ += w * J;
            mxx += w * ((P[0] * F00 + P[1] * F01) / J);
            myy += w * ((P[2] * F10 + P[3] * F11) / J);
            mxy += w * ((P[0] * F10 + P[1] * F11) / J);
            mzz += w * (P33 / J);
            mpsi += w * psi;
        }
        mJ /= area; mxx /= area; myy /= area; mxy /= area; mzz /= area; mpsi /= area;
        const double mises = sqrt(0.5 * ((mxx - myy) * (mxx - myy) + (myy - mzz) * (myy - mzz) +
                                         (mzz - mxx) * (mzz - mxx)) + 3.0 * mxy * mxy);
        if (t < ch.ne) {      // core elements (halo threads evaluate copies for their nodes only)
            const int eo = D.perm[e];
            const size_t ne0 = (size_t)D.n_el;
            eout[eo] = area;
            eout[ne0 + eo] = mJ;
            eout[2 * ne0 + eo] = mxx;
            eout[3 * ne0 + eo] = myy;
            eout[4 * ne0 + eo] = mxy;
            eout[5 * ne0 + eo] = mises;
            eout[6 * ne0 + eo] = mpsi;
        }
        val[0] = V2<double>{mJ, mxx};
        val[1] = V2<double>{myy, mxy};
        val[2] = V2<double>{mises, mpsi};
        val[3] = V2<double>{1.0, 0.0};
    }
    colour_accumulate4<NEN>(ch, t, l, val, L.acc, acc1, acc2, acc3);
    const size_t N = (size_t)D.n_nodes;
    V2<double>* wk = reinterpret_cast<V2<double>*>(work);
    for (int k = t; k < ch.nk; k += BS) {
        if (k < ch.nown) {
            const int n = (int)(D.lnode[ch.k0 + k] & DFEA_NODE_MASK);
            const double c = acc3[k].x;
            nout[n] = L.acc[k].x / c;
            nout[N + n] = L.acc[k].y / c;
            nout[2 * N + n] = acc1[k].x / c;
            nout[3 * N + n] = acc1[k].y / c;
            nout[4 * N + n] = acc2[k].x / c;
            nout[5 * N + n] = acc2[k].y / c;
        } else if (D.n_part) {
            const int slot = D.pslot[ch.pbase + (k - ch.nown)];
            wk[slot] = L.acc[k];
            wk[D.n_part + slot] = acc1[k];
            wk[2 * D.n_part + slot] = acc2[k];
            wk[3 * D.n_part + slot] = acc3[k];
        }
    }
}

__global__ __launch_bounds__(FBS) void k_fields_fin(const dfea_desc* __restrict__ tab, const int* __restrict__ items,
                                                   double* __restrict__ nout, const double* __restrict__ work) {
    const Item it = load_item(items);
    const dfea_desc& D = tab[it.s];
    const int i = it.b * FBS + threadIdx.x;
    if (i >= D.n_shared) return;
    const V2<double>* wk = reinterpret_cast<const V2<double>*>(work);
    const int j0 = D.sh_off[i], m = D.sh_off[i + 1] - j0;
    const V2<double> a0 = sum_slots<double>(wk, j0, m);
    const V2<double> a1 = sum_slots<double>(wk + D.n_part, j0, m);
    const V2<double> a2 = sum_slots<double>(wk + 2 * (size_t)D.n_part, j0, m);
    const V2<double> a3 = sum_slots<double>(wk + 3 * (size_t)D.n_part, j0, m);
    const size_t N = (size_t)D.n_nodes;
    const int n = (int)(D.sh_node[i] & DFEA_NODE_MASK);
    nout[n] = a0.x / a3.x;
    nout[N + n] = a0.y / a3.x;
    nout[2 * N + n] = a1.x / a3.x;
    nout[3 * N + n] = a1.y / a3.x;
    nout[4 * N + n] = a2.x / a3.x;
    nout[5 * N + n] = a2.y / a3.x;
}

// ---- adjoint: g[e] = -sum (m lam)_e . R^e_int(u; mat_p)  (`adjoint.py:75-106`) -------------------------------
template <int NEN, int NQ, int BS, bool LIB>
__global__ __launch_bounds__(BS) void k_adjoint(const dfea_desc* __restrict__ tab, const int* __restrict__ items,
                                               const double* __restrict__ u, const double* __restrict__ lam_in,
                                               double* __restrict__ gout) {
    __shared__ ChunkLDS<double> L;    // v: u, acc: masked lambda
    const Item it = load_item(items);
    const dfea_desc& D = tab[it.s];
    const ChunkView ch = chunk_view(D.chunks + it.b);
    const int t = threadIdx.x;
    for (int k = t; k < ch.nk; k += BS) {
        const uint32_t g = D.lnode[ch.k0 + k];
        const int n = (int)(g & DFEA_NODE_MASK);
        L.v[k] = reinterpret_cast<const V2<double>*>(u)[n];
        V2<double> lm = reinterpret_cast<const V2<double>*>(lam_in)[n];
        if (g & DFEA_FIXED_X) lm.x = 0.0;
        if (g & DFEA_FIXED_Y) lm.y = 0.0;
        L.acc[k] = lm;
    }
    __syncthreads();
    if (t >= ch.ne) return;                       // per-element output: core elements only
    const int ne = D.n_cols;
    const int e = ch.e0 + t;
    const dfea_mat::Mat M = dfea_mat::element_material<LIB>(D, 1.0);     // R^e_int is linear in theta_e
    V2<double> ue[NEN], le[NEN];
#pragma unroll
    for (int a = 0; a < NEN; ++a) {
        const int la = D.lelem[a * D.n_el + e];
        ue[a] = L.v[la];
        le[a] = L.acc[la];
    }
    double s = 0.0;
    for (int q = 0; q < NQ; ++q) {
        double d[NEN][2];
        double F00 = 1, F01 = 0, F10 = 0, F11 = 1;
#pragma unroll
        for (int a = 0; a < NEN; ++a) {
            d[a][0] = D.dN64[(size_t)((q * NEN + a) * 2) * ne + e];
            d[a][1] = D.dN64[(size_t)((q * NEN + a) * 2 + 1) * ne + e];
            F00 += ue[a].x * d[a][0]; F01 += ue[a].x * d[a][1];
            F10 += ue[a].y * d[a][0]; F11 += ue[a].y * d[a][1];
        }
        double P[4];
        piola_w<LIB>(M, F00, F01, F10, F11, D.w[(size_t)q * ne + e], P);
#pragma unroll
        for (int a = 0; a < NEN; ++a)
            s += le[a].x * (P[0] * d[a][0] + P[1] * d[a][1]) + le[a].y * (P[2] * d[a][0] + P[3] * d[a][1]);
    }
    gout[D.perm[e]] = -s;
}

// ======================================================================================================
// vector kernels (items {s, dof block})
// ======================================================================================================
__device__ __forceinline__ int ndof(const dfea_desc& D) { return 2 * D.n_nodes; }

// partial dot products of FP64 vectors -> red[dot region][blk]
__global__ __launch_bounds__(VBS) void k_dot_part(const dfea_desc* __restrict__ tab, const int* __restrict__ items,
                                                 const double* __restrict__ xin, const double* __restrict__ yin) {
    __shared__ double s_red[16];
    const Item it = load_item(items);
    const dfea_desc& D = tab[it.s];
    const double* x = xin ? xin : static_cast<const double*>(D.vec[DFEA_RES]);
    const double* y = yin ? yin : x;
    const int n = ndof(D);
    double s = 0.0;
    const int base = it.b * VBLK;
#pragma unroll
    for (int j = 0; j < VPT; ++j) {
        const int i = base + j * VBS + threadIdx.x;
        if (i < n) s += x[i] * y[i];
    }
    s = block_sum<VBS>(s, s_red);
    if (threadIdx.x == 0) vec_red(D)[2 * D.n_vblk + it.b] = s;
}

__global__ __launch_bounds__(VBS) void k_dot_final(const dfea_desc* __restrict__ tab, const int* __restrict__ items) {
    __shared__ double s_red[16];
    const Item it = load_item(items);
    const dfea_desc& D = tab[it.s];
    const double s = sum_partials<VBS>(vec_red(D) + 2 * D.n_vblk, D.n_vblk, s_red);
    if (threadIdx.x == 0) D.sc[DFEA_SC_NORM] = s;
}

// mode 0: out = where(fixed, lf uD, u);  mode 1: out = u + alpha dx;  mode 2: out = u (accept)
template <typename T>
__global__ __launch_bounds__(VBS) void k_axpy(const dfea_desc* __restrict__ tab, const int* __restrict__ items, int mode,
                                             const double* __restrict__ uin, const T* __restrict__ dxin,
                                             const double* __restrict__ uDin, const double* __restrict__ sarg,
                                             double* __restrict__ outp) {
    const Item it = load_item(items);
    const dfea_desc& D = tab[it.s];
    const int n = ndof(D);
    const double* u = uin ? uin : static_cast<const double*>(D.vec[mode == 2 ? DFEA_U : DFEA_UACC]);
    double* out = outp ? outp : static_cast<double*>(D.vec[mode == 2 ? DFEA_UACC : DFEA_U]);
    const double* uD = uDin ? uDin : static_cast<const double*>(D.vec[DFEA_UD]);
    const double a = mode == 2 ? 0.0 : sarg[it.s];
    const T* dx = dxin ? dxin : static_cast<const T*>(D.vec[DFEA_X]);
    const int base = it.b * VBLK;
#pragma unroll
    for (int j = 0; j < VPT; ++j) {
        const int i = base + j * VBS + threadIdx.x;
        if (i >= n) continue;
        if (mode == 0) {
            const bool fixed = (D.nfix[i >> 1] >> (i & 1)) & 1;
            out[i] = fixed ? a * uD[i] : u[i];
        } else if (mode == 1) {
            out[i] = u[i] + a * (double)dx[i];
        } else {
            out[i] = u[i];
        }
    }
}

// ======================================================================================================
// PCG (`newton_krylov.py:28-88`) in one kernel per iteration.
//
// The recurrences are those of Jacobi-PCG in the Chronopoulos-Gear arrangement, which yields the same iterates
// as the textbook form in exact arithmetic but needs one reduction phase per iteration: with u = M r and
// w = A u, gamma = r.u and delta = w.u are summed together, and
//     beta_k = gamma_k / gamma_{k-1},  p.Ap = delta_k - beta_k gamma_k / alpha_{k-1},  alpha_k = gamma_k / p.Ap,
//     p_k = u_k + beta_k p_{k-1},  s_k = w_k + beta_k s_{k-1} (= A p_k),
//     x_{k+1} = x_k + alpha_k p_k,  r_{k+1} = r_k - alpha_k s_k,  u_{k+1} = M r_{k+1},  w_{k+1} = A u_{k+1}.
// Kernel K_k (k = -1 computes w_0 = A u_0 only) opens with the scalar step from the previous kernel's chunk
// partials, summed in a fixed order by every workgroup (all take the same decision): convergence of r_k,
// the iteration cap, curvature p.Ap <= 0 (x = z = u_0 at k = 0) -- the reference's tests, in its order.  It
// then updates every node of its chunk (shared nodes redundantly, with w of shared nodes summed from the
// previous kernel's partials in slot order; owned nodes and a shared node's first-slot chunk write), and
// applies A to u_{k+1} in LDS.  gamma and r.r come from the nodes the chunk writes; delta of a shared node is
// linear in its partials (u . this chunk's partial, fixed components u^2 from the first-slot chunk).
//
// Kernels are launched in pairs with parity arguments 0, 1: K_k has parity (k + 1) mod 2, reads the state
// buffers of parity 1 - par (x_k, r_k, p_{k-1}, s_{k-1}, w_k), the partials and the scalar set of parity
// 1 - par / par, and writes those of the other parity, so no workgroup reads what another workgroup of the
// same launch writes.  The solution ends in vec[DFEA_X].
// ======================================================================================================
enum { PS_ST = 0, PS_K = 1, PS_GAM = 2, PS_ALPHA = 3 };    // fields of a parity scalar set

// the PCG state buffers of parity q (0: the slot, 1: its twin)
template <typename T>
struct PcgSet {
    V2<T>* x;
    V2<T>* r;
    V2<T>* w;
    V2<T>* p;
    V2<T>* s;
};

template <typename T>
__device__ __forceinline__ PcgSet<T> pcg_set(const dfea_desc& D, int q) {
    PcgSet<T> v;
    v.x = static_cast<V2<T>*>(D.vec[q ? DFEA_X2 : DFEA_X]);
    v.r = static_cast<V2<T>*>(D.vec[q ? DFEA_R2 : DFEA_R]);
    v.w = static_cast<V2<T>*>(D.vec[q ? DFEA_W2 : DFEA_W]);
    v.p = static_cast<V2<T>*>(D.vec[q ? DFEA_P2 : DFEA_P]);
    v.s = static_cast<V2<T>*>(D.vec[q ? DFEA_S2 : DFEA_S]);
    return v;
}

// PCG start: x = 0, r = b, p = s = 0 (parity-0 buffers) and r = b in the parity-1 buffer (the first kernel, of
// parity 0, reads r_0 there like every kernel reads the buffers of the other parity); partials of b.b
template <typename T>
__global__ __launch_bounds__(VBS) void k_pcg_init(const dfea_desc* __restrict__ tab, const int* __restrict__ items,
                                                 int neg_res) {
    __shared__ double s_red[16];
    const Item it = load_item(items);
    const dfea_desc& D = tab[it.s];
    const int n = ndof(D);
    T* x = static_cast<T*>(D.vec[DFEA_X]);
    T* r = static_cast<T*>(D.vec[DFEA_R]);
    T* r2 = static_cast<T*>(D.vec[DFEA_R2]);
    T* p = static_cast<T*>(D.vec[DFEA_P]);
    T* sv = static_cast<T*>(D.vec[DFEA_S]);
    const T* bT = static_cast<const T*>(D.vec[DFEA_B]);
    const double* res = static_cast<const double*>(D.vec[DFEA_RES]);
    double bb = 0.0;
    const int base = it.b * VBLK;
#pragma unroll
    for (int j = 0; j < VPT; ++j) {
        const int i = base + j * VBS + threadIdx.x;
        if (i >= n) continue;
        const T b = neg_res ? (T)(-res[i]) : bT[i];
        x[i] = 0;
        r[i] = b;
        r2[i] = b;
        p[i] = 0;
        sv[i] = 0;
        bb += (double)b * (double)b;
    }
    bb = block_sum<VBS>(bb, s_red);
    if (threadIdx.x == 0) vec_red(D)[it.b] = bb;
}

// ||b||^2 and the scalar block: both parity sets start the solve (k = -1: the first kernel computes w_0);
// set: the tolerance and the cap come from the arguments (else they are already in the scalar block)
__global__ __launch_bounds__(VBS) void k_pcg_init_final(const dfea_desc* __restrict__ tab, const int* __restrict__ items,
                                                       int set, double rtol, double maxit) {
    __shared__ double s_red[16];
    const Item it = load_item(items);
    const dfea_desc& D = tab[it.s];
    const double bb = sum_partials<VBS>(vec_red(D), D.n_vblk, s_red);
    if (threadIdx.x == 0) {
        double* sc = D.sc;
        // nothing to do when the right-hand side is zero: x = 0 after 0 iterations
        const double st = sqrt(bb) == 0.0 ? (double)DFEA_PCG_CONVERGED : (double)DFEA_PCG_RUNNING;
        if (set) {
            sc[DFEA_SC_RTOL] = rtol;
            sc[DFEA_SC_MAXIT] = maxit;
        }
        sc[DFEA_SC_BB] = bb;
        sc[DFEA_SC_RR] = bb;
        sc[DFEA_SC_RZ] = 0.0;
        sc[DFEA_SC_PAP] = 0.0;
        sc[DFEA_SC_ALPHA] = 0.0;
        sc[DFEA_SC_BETA] = 0.0;
        sc[DFEA_SC_K] = 0.0;
        sc[DFEA_SC_ITERS] = 0.0;
        sc[DFEA_SC_STATUS] = st;
        for (int par = 0; par < 2; ++par) {
            double* S = sc + (par ? DFEA_SC_S1 : DFEA_SC_S0);
            S[PS_ST] = st;
            S[PS_K] = -1.0;
            S[PS_GAM] = 0.0;
            S[PS_ALPHA] = 0.0;
        }
    }
}

// the state of local node kk of a chunk for the PCG update: x_k, r_k, p_{k-1}, s_{k-1}, w_k (w of a shared
// node is the slot-order sum of the previous kernel's partials: node_state loads the first SUM_UNROLL of them
// into pv, node_w adds them), the preconditioner (d: the Jacobi diagonal, or the inverse block's diagonal with
// its off-diagonal entry o), fixed bits, and whether this chunk writes the node (owned nodes; a shared node's
// first-slot chunk; halo plans: nodes past nown are owned by another chunk and only read)
template <typename T>
struct NodeState {
    V2<T> x, r, p, s, w, d;
    V2<T> pv[SUM_UNROLL];
    T o;
    int n, j0, cnt;
    unsigned char b;
    bool writer, sh;
};

// u = M r: r / diag (point Jacobi) or the inverse nodal block times r (block Jacobi, dgo != nullptr)
template <typename T>
__device__ __forceinline__ V2<T> precond_apply(V2<T> r, V2<T> d, T o, bool blk) {
    if (blk) return V2<T>{d.x * r.x + o * r.y, o * r.x + d.y * r.y};
    return V2<T>{r.x / d.x, r.y / d.y};
}

// the state of local node kk whose node id g, slot range rg and slot sl (shared entries) are already loaded;
// DEFER: a shared node's partials are left in pv for node_w (else w is summed here)
template <typename T, bool DEFER>
__device__ __forceinline__ NodeState<T> node_state(const dfea_desc& D, const ChunkView& ch, int kk, uint32_t g,
                                                   int2 rg, int sl, bool upd, const PcgSet<T>& Rd, const V2<T>* dg,
                                                   const T* dgo, const V2<T>* part_rd) {
    NodeState<T> ns;
    ns.n = (int)(g & DFEA_NODE_MASK);
    ns.b = (unsigned char)(g >> 30);
    const bool shared = kk >= ch.nown && D.n_part;
    ns.writer = shared ? rg.x == sl : kk < ch.nown;
    ns.d = gload2<T>(dg, (uint32_t)ns.n);
    ns.o = dgo ? gload<false>(dgo, (uint32_t)ns.n) : (T)0;
    ns.r = gload2<T>(Rd.r, (uint32_t)ns.n);
    ns.sh = DEFER && upd && shared;
    // DEFER (early-loaded states): requested whatever k is (the buffers of the other parity always exist; the
    // first kernel, k < 0, drops x, p, s, w), so the loads need not wait for the iteration counter
    if (DEFER || upd) {
        if (shared && !DEFER) {
            ns.w = sum_slots<T>(part_rd, rg.x, rg.y);
        } else if (shared) {
            ns.j0 = rg.x;
            ns.cnt = rg.y;
#pragma unroll
            for (int q = 0; q < SUM_UNROLL; ++q)
                if (q < rg.y) ns.pv[q] = part_rd[rg.x + q];
        } else {
            ns.w = gload2<T>(Rd.w, (uint32_t)ns.n);
        }
        ns.x = gload2<T>(Rd.x, (uint32_t)ns.n);
        ns.p = gload2<T>(Rd.p, (uint32_t)ns.n);
        ns.s = gload2<T>(Rd.s, (uint32_t)ns.n);
    }
    return ns;
}

// w of a shared node: its partials added in slot order, as sum_slots adds them (same operations, same bits);
// called where w is used, so the loads of an early-loaded state stay in flight across the scalar step
template <typename T>
__device__ __forceinline__ void node_w(NodeState<T>& ns, const V2<T>* part_rd) {
    if (!ns.sh) return;
    V2<T> acc{0, 0};
#pragma unroll
    for (int q = 0; q < SUM_UNROLL; ++q) {
        if (q < ns.cnt) {
            acc.x += ns.pv[q].x;
            acc.y += ns.pv[q].y;
        }
    }
    for (int j = ns.j0 + SUM_UNROLL; j < ns.j0 + ns.cnt; ++j) {
        const V2<T> pj = part_rd[j];
        acc.x += pj.x;
        acc.y += pj.y;
    }
    ns.w = acc;
}

// one PCG iteration (see above); par = launch parity; MINW: minimum waves per SIMD for the register budget.
// Loads return in the order they are issued (waiting for one waits for those before it, not for those after),
// so the kernel issues everything that does not depend on the scalars first, in the order it is needed: the
// previous kernel's chunk partials (scalar step), the node ids and slot tables of the thread's nodes, the
// element's local node indices, the first PF Gauss points.  EARLY: the state of every node of the thread is also
// loaded before the scalar step, so its loads overlap the reduction (and the callers pick PF = all the group's
// Gauss points); it costs registers (waves per SIMD), which matters only when the launch has more workgroups
// than the GPU has CUs.  The scalars the step reads (tolerance, cap, the previous iteration's gamma and
// alpha) are requested at the start, before the first barrier.
template <int NEN, int NQ, int BS, typename T, int PF = 2, int MINW = 1, bool EARLY = false,
          int CEX = default_ce<NEN>(), bool NT = false>
__global__ __launch_bounds__(BS, MINW) void k_pcg(const dfea_desc* __restrict__ tab, const int* __restrict__ items, int par) {
    static_assert(3 * (BS / 64) <= 16, "block_sum3 needs 3 slots per wave in ChunkLDS::red");
    // partials plans (chunks of exactly CE elements) sum nodes through the inverse map, halo plans by colour
    constexpr bool INV = BS % Groups<NEN, BS, CEX>::CE == 0;
    __shared__ std::conditional_t<INV, ProductLDS<T, CEX * NEN>, ChunkLDS<T>> L;
    __shared__ GroupLDS<NEN, BS, T, CEX> GL;
    const Item it = load_item(items);
    const dfea_desc& D = tab[it.s];
    double* sc = D.sc;
    const double* S = sc + (par ? DFEA_SC_S1 : DFEA_SC_S0);
    // the status is tested once the chunk record and the chunk partials are requested (their loads are valid
    // whatever the status), so a running kernel does not wait for the status before issuing them
    const double st0 = S[PS_ST];
    const double k = S[PS_K];
    // the step's scalars, requested now (a load behind the reduction's barrier would wait a round trip there)
    const double s_rtol = sc[DFEA_SC_RTOL], s_bb = sc[DFEA_SC_BB], s_maxit = sc[DFEA_SC_MAXIT];
    const double s_gam = S[PS_GAM], s_alpha = S[PS_ALPHA];
    const ChunkView ch = chunk_view(D.chunks + it.b);
    const int t = threadIdx.x;
    const bool lead = it.b == 0 && t == 0;
    const int ne = D.n_cols;
    const T* dN = std::is_same<T, double>::value ? (const T*)D.dN64 : (const T*)D.dN32;
    const T* C = std::is_same<T, double>::value ? (const T*)D.C64 : (const T*)D.C32;
    const V2<T>* dg = reinterpret_cast<const V2<T>*>(std::is_same<T, double>::value ? (const T*)D.diag64
                                                                                       : (const T*)D.diag32);
    const bool blk = D.precond == DFEA_PRECOND_BLOCK;
    const T* dgo = blk ? reinterpret_cast<const T*>(dg + D.n_nodes) : nullptr;
    // state buffers read (x_k, r_k, p_{k-1}, s_{k-1}, w_k; the first kernel, k < 0, only r_0, which dfea_pcg_init
    // writes to both parities): EARLY launches take parity 1 - par for every k, so the buffer addresses do not
    // wait for the iteration counter; larger launches (measured faster so) keep parity par for k < 0
    const int qr = EARLY || k >= 0.0 ? 1 - par : par;
    const PcgSet<T> Rd = pcg_set<T>(D, qr);
    const PcgSet<T> Wr = pcg_set<T>(D, par);   // state buffers written
    // the shared-node partials of parity q: FP64 pairs q n_part ..; FP32 solves keep theirs in the buffer's second
    // half (float2 offset (2 + q) n_part), clear of the FP64 pairs the residual kernels write from its start, so
    // a residual of the same solve may run between two rounds of PCG kernels
    constexpr size_t PB = sizeof(T) == 4 ? 2 : 0;
    const V2<T>* part_rd = reinterpret_cast<const V2<T>*>(D.part) + (PB + 1 - par) * (size_t)D.n_part;
    V2<T>* part_wr = reinterpret_cast<V2<T>*>(D.part) + (PB + par) * (size_t)D.n_part;
    const int2* prg = reinterpret_cast<const int2*>(D.prange);
    constexpr int CE = Groups<NEN, BS, CEX>::CE;
    constexpr int NG = (MAXLN + BS - 1) / BS;     // local nodes per thread, at most
    constexpr int NRD = 3;                        // chunk partials per thread loaded ahead of the scalar step

    // 1a. the previous kernel's chunk partials {r.u, w.u, r.r} (index clamped: extra threads read chunk 0 and
    // drop it; the first kernel, k < 0, reads and drops them)
    const double* rd = pcg_red(D, 1 - par);
    const bool presum = D.n_chunks > PCG_PRESUM_CHUNKS;      // the totals come from k_pcg_reduce
    double pr[NRD][3];
    double tot[3] = {0.0, 0.0, 0.0};
    auto load_partials = [&]() {
        if (presum) {
            const double* tt = pcg_tot(D, 1 - par);
            tot[0] = tt[0];
            tot[1] = tt[1];
            tot[2] = tt[2];
        } else {
#pragma unroll
            for (int j = 0; j < NRD; ++j) {
                const int i = t + j * BS;
                const uint32_t ic = (uint32_t)(i < D.n_chunks ? i : 0);
#pragma unroll
                for (int m = 0; m < 3; ++m) pr[j][m] = gload<false>(rd, 3u * ic + m);
            }
        }
    };
    // 1b. node ids, and for shared entries the slot range of the node and the chunk's own slot
    uint32_t gid[NG];
    int2 rng[NG];
    int slt[NG];
    auto load_nodes = [&]() {
#pragma unroll
        for (int i = 0; i < NG; ++i) {
            const int kk = t + i * BS;
            const int kc = kk < ch.nk ? kk : ch.nk - 1;
            gid[i] = gload<false>(D.lnode, INV ? lentry(D, it.b, kk) : (uint32_t)(ch.k0 + kc));
            rng[i] = make_int2(0, 0);
            slt[i] = 0;
            if (D.n_part) {
                // EARLY (small launches, latency-bound): by the fixed stride, without waiting for the chunk record;
                // else from the record, so the entries of owned nodes share one cache line (less traffic)
                const uint32_t ent = INV && EARLY ? lentry(D, it.b, kk)
                                                  : (uint32_t)(kc >= ch.nown ? ch.pbase + (kc - ch.nown) : 0);
                const V2<int> r2 = gload2<int>(reinterpret_cast<const V2<int>*>(prg), ent);
                rng[i] = make_int2(r2.x, r2.y);
                slt[i] = gload<false>(D.pslot, ent);
            }
        }
    };
    // EARLY: the node ids first -- the node states wait for them, and the partials (written by every chunk of
    // the previous kernel, so fetched from beyond the XCD's L2) would hold them back (loads return in order)
    if constexpr (EARLY) {
        load_nodes();
        load_partials();
    } else {
        load_partials();
        load_nodes();
    }
    if (st0 != DFEA_PCG_RUNNING) return;
    // 1c. element (partials plans: tail threads of the last chunk read the last element and drop it) and its
    // first PF Gauss points
    int l[NEN];
    int e;
    bool has_el;
    InvRegs<NEN, BS, CEX> inv;
    if constexpr (BS % CE == 0) {
        const int tc = t % CE;
        has_el = tc < ch.ne;
        e = ch.e0 + (has_el ? tc : ch.ne - 1);
#pragma unroll
        for (int a = 0; a < NEN; ++a) l[a] = gload<false>(D.lelem, (uint32_t)(a * D.n_el + e));
        inv_load<NEN, BS, CEX, true>(D, ch, it.b, t, inv);
    } else {
        has_el = thread_element<NEN>(D, ch, t, e, l);
    }
    constexpr int QG = GaussSplit<NQ, Groups<NEN, BS, CEX>::G>::QG;
    T d[QG][NEN][2];
    T c[QG][10];
    __builtin_amdgcn_sched_barrier(0);
    prefetch_gauss_points<NEN, NQ, BS, PF, T, NT, CEX>(dN, C, ne, e, t, d, c);
    __builtin_amdgcn_sched_barrier(0);

    // 1d. the node states (independent of the scalars)
    // EARLY: the first NGE node slots of the thread (chunks of up to 384 nodes: every structured chunk)
    constexpr int NGE = (384 + BS - 1) / BS < NG ? (384 + BS - 1) / BS : NG;
    NodeState<T> nse[EARLY ? NGE : 1];
    if constexpr (EARLY) {
#pragma unroll
        for (int i = 0; i < NGE; ++i) {
            const int kk = t + i * BS;
            if (kk < ch.nk) nse[i] = node_state<T, true>(D, ch, kk, gid[i], rng[i], slt[i], k >= 0.0, Rd, dg, dgo, part_rd);
        }
    }

    // 2. scalar step
    double alpha = 0.0, beta = 0.0, gam = 0.0, den = 0.0, rr = 0.0;
    if (k >= 0.0) {
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        if (presum) {
            a0 = tot[0];
            a1 = tot[1];
            a2 = tot[2];
        } else {
#pragma unroll
            for (int j = 0; j < NRD; ++j) {
                if (t + j * BS < D.n_chunks) {
                    a0 += pr[j][0];
                    a1 += pr[j][1];
                    a2 += pr[j][2];
                }
            }
            for (int i = t + NRD * BS; i < D.n_chunks; i += BS) {
                a0 += rd[3 * i];
                a1 += rd[3 * i + 1];
                a2 += rd[3 * i + 2];
            }
            block_sum3<BS, true>(a0, a1, a2, L.red);     // L.red is next written after the gather barrier
        }
        gam = a0;
        const double delta = a1;
        rr = a2;
        double st = DFEA_PCG_RUNNING, iters = 0.0;
        if (k >= 1.0 && sqrt(rr) <= s_rtol * sqrt(s_bb)) {
            st = DFEA_PCG_CONVERGED;
            iters = k;
        } else if (k >= s_maxit) {
            st = DFEA_PCG_MAXITER;
            iters = s_maxit;
        } else {
            beta = k >= 1.0 ? gam / s_gam : 0.0;
            den = k >= 1.0 ? delta - beta * gam / s_alpha : delta;
            if (den <= 0.0) {      // `newton_krylov.py:70-71`: NaN runs on, as in the reference
                st = DFEA_PCG_CURVATURE;
                iters = k;
            } else {
                alpha = gam / den;
            }
        }
        if (st != DFEA_PCG_RUNNING) {
            if (lead) {
                double* O = sc + (par ? DFEA_SC_S0 : DFEA_SC_S1);
                O[PS_ST] = st;
                O[PS_K] = k;
                sc[DFEA_SC_STATUS] = st;
                sc[DFEA_SC_ITERS] = iters;
                sc[DFEA_SC_K] = k;
                sc[DFEA_SC_RR] = rr;
                sc[DFEA_SC_PAP] = den;
            }
            // the answer into vec[DFEA_X]: x_k lives in buffer k mod 2; curvature at k = 0 answers z = u_0
            const bool at0 = st == DFEA_PCG_CURVATURE && k == 0.0;
            if (at0 || qr == 1) {
                V2<T>* x0 = static_cast<V2<T>*>(D.vec[DFEA_X]);
                for (int kk = t; kk < ch.nk; kk += BS) {
                    const int n = (int)(D.lnode[ch.k0 + kk] & DFEA_NODE_MASK);
                    if (kk >= ch.nown && (!D.n_part || prg[ch.pbase + (kk - ch.nown)].x != D.pslot[ch.pbase + (kk - ch.nown)]))
                        continue;
                    if (at0) {
                        x0[n] = precond_apply<T>(Rd.r[n], dg[n], blk ? dgo[n] : (T)0, blk);
                    } else {
                        x0[n] = Rd.x[n];
                    }
                }
            }
            return;
        }
    }

    // 3. update of the chunk's nodes, u_{k+1} into LDS
    double g_part = 0.0, r_part = 0.0;
#pragma unroll
    for (int i = 0; i < NG; ++i) {
        const int kk = t + i * BS;
        if (kk >= ch.nk) break;
        NodeState<T> ns;
        if (EARLY && i < NGE) ns = nse[EARLY ? (i < NGE ? i : 0) : 0];
        else ns = node_state<T, false>(D, ch, kk, gid[i], rng[i], slt[i], k >= 0.0, Rd, dg, dgo, part_rd);
        node_w<T>(ns, part_rd);
        const int n = ns.n;
        const unsigned char b = ns.b;
        V2<T> rn = ns.r;
        const V2<T> dv = ns.d;
        if (k >= 0.0) {
            V2<T> w = ns.w;
            const V2<T> u = precond_apply<T>(rn, dv, ns.o, blk);
            if (b & 1) w.x = u.x;       // fixed dofs: A acts as the identity
            if (b & 2) w.y = u.y;
            V2<T> pn, sn, xn;
            pn.x = (T)((double)u.x + beta * (double)ns.p.x);
            pn.y = (T)((double)u.y + beta * (double)ns.p.y);
            sn.x = (T)((double)w.x + beta * (double)ns.s.x);
            sn.y = (T)((double)w.y + beta * (double)ns.s.y);
            xn.x = (T)((double)ns.x.x + alpha * (double)pn.x);
            xn.y = (T)((double)ns.x.y + alpha * (double)pn.y);
            rn.x = (T)((double)rn.x - alpha * (double)sn.x);
            rn.y = (T)((double)rn.y - alpha * (double)sn.y);
            if (ns.writer) {
                Wr.p[n] = pn;
                Wr.s[n] = sn;
                Wr.x[n] = xn;
                Wr.r[n] = rn;
            }
        }
        const V2<T> un = precond_apply<T>(rn, dv, ns.o, blk);
        if (ns.writer) {
            g_part += (double)rn.x * (double)un.x + (double)rn.y * (double)un.y;
            r_part += (double)rn.x * (double)rn.x + (double)rn.y * (double)rn.y;
        }
        L.v[kk] = un;
        if constexpr (!INV) L.acc[kk] = V2<T>{0, 0};
        L.bits[kk] = b;
    }
    if constexpr (INV) inv_store<NEN, BS, T, CEX>(ch, t, inv, L);
    __syncthreads();

    // 4. w_{k+1} = A u_{k+1}
    element_product<NEN, NQ, BS, PF, T, NT, false, CEX>(ch, t, has_el, l, dN, C, ne, e, d, c, L, GL);

    double d_part = 0.0;
#pragma unroll
    for (int i = 0; i < NG; ++i) {
        const int kk = t + i * BS;
        if (kk >= ch.nk) break;
        const V2<T> u = L.v[kk];
        V2<T> acc;
        if constexpr (INV) acc = node_sum<T>(L, kk);
        else acc = L.acc[kk];
        const unsigned char b = L.bits[kk];
        if (kk < ch.nown) {
            V2<T> out;
            out.x = (b & 1) ? u.x : acc.x;
            out.y = (b & 2) ? u.y : acc.y;
            Wr.w[(int)(gid[i] & DFEA_NODE_MASK)] = out;
            d_part += (double)u.x * (double)out.x + (double)u.y * (double)out.y;
        } else if (D.n_part) {
            const int slot = slt[i];
            part_wr[slot] = acc;
            if (!(b & 1)) d_part += (double)u.x * (double)acc.x;
            if (!(b & 2)) d_part += (double)u.y * (double)acc.y;
            if (b) {     // fixed components (w = u there) once, by the chunk holding the node's first slot
                if (slot == rng[i].x) {
                    if (b & 1) d_part += (double)u.x * (double)u.x;
                    if (b & 2) d_part += (double)u.y * (double)u.y;
                }
            }
        }
    }
    block_sum3<BS, true>(g_part, d_part, r_part, L.red);     // last use of L.red
    if (t == 0) {
        double* wr = pcg_red(D, par) + 3 * it.b;
        wr[0] = g_part;
        wr[1] = d_part;
        wr[2] = r_part;
    }
    if (lead) {
        double* O = sc + (par ? DFEA_SC_S0 : DFEA_SC_S1);
        O[PS_ST] = DFEA_PCG_RUNNING;
        O[PS_K] = k + 1.0;
        O[PS_GAM] = gam;
        O[PS_ALPHA] = alpha;
        if (k >= 0.0) {
            sc[DFEA_SC_K] = k + 1.0;
            sc[DFEA_SC_RZ] = gam;
            sc[DFEA_SC_PAP] = den;
            sc[DFEA_SC_ALPHA] = alpha;
            sc[DFEA_SC_BETA] = beta;
            sc[DFEA_SC_RR] = rr;
        }
    }
}

// the totals {r.u, w.u, r.r} of the chunk partials a PCG kernel of parity par wrote, for every solve of more than
// PCG_PRESUM_CHUNKS chunks: one workgroup per solve (the solves are found as their block-0 items), the same fixed
// order of additions every time (thread-strided sums, then the workgroup tree)
__global__ __launch_bounds__(VBS) void k_pcg_reduce(const dfea_desc* __restrict__ tab, const int* __restrict__ items,
                                                   int n_items, int par) {
    __shared__ double s_red[3 * (VBS / 64)];
    for (int j = blockIdx.x; j < n_items; j += gridDim.x) {
        const int sidx = items ? items[2 * j] : 0, b = items ? items[2 * j + 1] : j;
        if (b != 0) continue;
        const dfea_desc& D = tab[sidx];
        if (D.n_chunks <= PCG_PRESUM_CHUNKS) continue;
        const double* rd = pcg_red(D, par);
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int i = threadIdx.x; i < D.n_chunks; i += VBS) {
            a0 += rd[3 * i];
            a1 += rd[3 * i + 1];
            a2 += rd[3 * i + 2];
        }
        block_sum3<VBS>(a0, a1, a2, s_red);
        if (threadIdx.x == 0) {
            double* tt = pcg_tot(D, par);
            tt[0] = a0;
            tt[1] = a1;
            tt[2] = a2;
        }
    }
}

// ======================================================================================================
// host launch helpers
// ======================================================================================================
inline hipStream_t S(void* s) { return static_cast<hipStream_t>(s); }

template <typename K, typename... A>
int launch(K kern, int grid, int block, void* stream, const char* name, A... args) {
    if (grid <= 0) return 0;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, S(stream), args...);
    return hip_status(name);
}

// element kind of a launch: nen, with DFEA_HALO set for halo plans (DFEA_MATLIB selects the instantiations
// with the whole material library, DFEA_TAIL marks tail-balanced linear partials plans, DFEA_WIDE the PCG
// workgroups of small quadratic meshes; none of them is part of the kind)
enum class Kind { T3, Q4, Q4H, Q4S, T6, Q8, Q9, Bad };
Kind kind_of(int k) {
    k &= ~DFEA_MATLIB;
    if (k & DFEA_WIDE) {
        k &= ~DFEA_WIDE;
        if (k != 6 && k != 8 && k != 9) return Kind::Bad;
    }
    if (k & DFEA_TAIL) {
        k &= ~DFEA_TAIL;
        if (k != 3 && k != 4) return Kind::Bad;
    }
    switch (k) {
        case 3: return Kind::T3;
        case 4: return Kind::Q4;
        case 4 | DFEA_HALO: return Kind::Q4H;
        case 4 | DFEA_SMALL: return Kind::Q4S;
        case 6: return Kind::T6;
        case 8: return Kind::Q8;
        case 9: return Kind::Q9;
        default: return Kind::Bad;
    }
}

// call f.template run<NEN, BS, LIB>() for the kind (one instantiation per element type, plan mode and, for
// the kernels that evaluate the material, library or neo-Hookean-only code)
template <bool LIB, typename F>
int dispatch_lib(int kind, F f) {
    switch (kind_of(kind)) {
        case Kind::T3: return f.template run<3, BS_LIN, LIB>();
        case Kind::Q4: return f.template run<4, BS_LIN, LIB>();
        case Kind::Q4H: return f.template run<4, BS_Q4H, LIB>();
        case Kind::Q4S: return f.template run<4, BS_QUAD, LIB>();
        case Kind::T6: return f.template run<6, BS_QUAD, LIB>();
        case Kind::Q8: return f.template run<8, BS_QUAD, LIB>();
        case Kind::Q9: return f.template run<9, BS_QUAD, LIB>();
        default:
            return fail(DFEA_EINVAL, "unknown element kind (expected 3, 4, 6, 8, 9, 4 | DFEA_HALO, 4 | DFEA_SMALL, "
                                     "3 | DFEA_TAIL or 4 | DFEA_TAIL, optionally | DFEA_MATLIB)");
    }
}

template <typename F>
int dispatch(int kind, F f) {
    return (kind & DFEA_MATLIB) ? dispatch_lib<true>(kind, f) : dispatch_lib<false>(kind, f);
}

// PCG kernels in pairs of parities 0, 1 (n_iter rounded up to even).  Quadratic-element chunks (64 elements,
// up to 289 nodes) run with bs_groups threads: Gauss-point groups per element, node updates on all lanes.
// WIDE (DFEA_WIDE meshes of quadratic elements): workgroups of bs_groups_wide threads, whatever the launch size,
// so a solve's arithmetic does not depend on the batch it runs in
template <int NEN, int BS, typename T, bool WIDE = false>
int pcg_iter(const dfea_desc* table, const int32_t* items, int n_items, int n_iter, void* stream) {
    // small-chunk quad4 plans (dispatched with the quadratic elements' chunk workgroup): 64-element chunks,
    // 4 threads per element (one Gauss point each)
    constexpr bool SMALL = NEN == 4 && BS == BS_QUAD;
    constexpr int CEX = SMALL ? CHUNK_QUAD : default_ce<NEN>();
    constexpr int BSP = quadratic<NEN>() ? (WIDE ? bs_groups_wide<T>() : bs_groups<T>()) : SMALL ? BS_Q4S : BS;
    const bool early = n_items <= PCG_EARLY_ITEMS;
    // element data beyond what the Infinity Cache holds next to the state vectors: streamed with nontemporal loads
    const bool nt = (double)n_items * CEX * nq_of<NEN>() * (10 + 2 * NEN) * sizeof(T) > PCG_NT_BYTES;
    for (int i = 0; i < n_iter + (n_iter & 1); ++i) {
        // small launches (one workgroup per CU or fewer): every node state and, for linear elements, every Gauss
        // point requested before the scalar step
        constexpr int PF_EARLY = NEN <= 4 && BSP % CHUNK_LIN == 0 ? nq_of<NEN>() : 2;
        constexpr int MW = NEN <= 4 && !SMALL ? 3 : 1;
        const int rc = early ? launch(k_pcg<NEN, nq_of<NEN>(), BSP, T, PF_EARLY, 1, true, CEX>, n_items, BSP, stream,
                                      "k_pcg", table, items, i & 1)
                     : nt ? launch(k_pcg<NEN, nq_of<NEN>(), BSP, T, 2, MW, false, CEX, true>, n_items, BSP, stream,
                                   "k_pcg", table, items, i & 1)
                          : launch(k_pcg<NEN, nq_of<NEN>(), BSP, T, 2, MW, false, CEX>, n_items, BSP, stream,
                                   "k_pcg", table, items, i & 1);
        if (rc) return rc;
        if (n_items > PCG_PRESUM_CHUNKS) {      // a solve of the launch may have more chunks: sum its partials
            const int grid = n_items < PCG_REDUCE_GRID ? n_items : PCG_REDUCE_GRID;
            const int rr = launch(k_pcg_reduce, grid, VBS, stream, "k_pcg_reduce", table, items, n_items, i & 1);
            if (rr) return rr;
        }
    }
    return 0;
}


// launchers of the C ABI, one instantiation per element kind
struct Launch_geometry {
    const dfea_desc* t; const int32_t* it; int n; void* st;
    template <int NEN, int BS, bool LIB> int run() {
        return launch(k_geometry<NEN, nq_of<NEN>(), BS>, n, BS, st, "k_geometry", t, it);
    }
};

struct Launch_moduli {
    const dfea_desc* t; const int32_t* it; int n; const int32_t* fi; int nf; const double* u; int prec; void* st;
    template <int NEN, int BS, bool LIB> int run() {
        int rc = prec == DFEA_FP64 ? launch(k_moduli<NEN, nq_of<NEN>(), BS, double, LIB>, n, BS, st, "k_moduli", t, it, u)
                                   : launch(k_moduli<NEN, nq_of<NEN>(), BS, float, LIB>, n, BS, st, "k_moduli", t, it, u);
        if (rc) return rc;
        return prec == DFEA_FP64 ? launch(k_moduli_fin<double>, nf, FBS, st, "k_moduli_fin", t, fi)
                                 : launch(k_moduli_fin<float>, nf, FBS, st, "k_moduli_fin", t, fi);
    }
};

struct Launch_residual {
    const dfea_desc* t; const int32_t* it; int n; const int32_t* fi; int nf; const double* u; double* R;
    double lf; void* st;
    template <int NEN, int BS, bool LIB> int run() {
        int rc = launch(k_residual<NEN, nq_of<NEN>(), BS, LIB>, n, BS, st, "k_residual", t, it, u, R, lf);
        if (rc) return rc;
        return launch(k_residual_fin, nf, FBS, st, "k_residual_fin", t, fi, R, lf);
    }
};

struct Launch_reactions {
    const dfea_desc* t; const int32_t* it; int n; const int32_t* si; int ns; const double* u; double* out; void* st;
    template <int NEN, int BS, bool LIB> int run() {
        int rc = launch(k_residual<NEN, nq_of<NEN>(), BS, LIB, true>, n, BS, st, "k_residual<react>", t, it, u, (double*)nullptr, 1.0);
        if (rc) return rc;
        return launch(k_reactions_final, ns, VBS, st, "k_reactions_final", t, si, out);
    }
};

struct Launch_fields {
    const dfea_desc* t; const int32_t* it; int n; const int32_t* fi; int nf; const double* u; double* eout;
    double* nout; double* work; void* st;
    template <int NEN, int BS, bool LIB> int run() {
        int rc = launch(k_fields<NEN, nq_of<NEN>(), BS, LIB>, n, BS, st, "k_fields", t, it, u, eout, nout, work);
        if (rc) return rc;
        return launch(k_fields_fin, nf, FBS, st, "k_fields_fin", t, fi, nout, (const double*)work);
    }
};

struct Launch_apply {
    const dfea_desc* t; const int32_t* it; int n; const int32_t* fi; int nf; const void* v; void* y; int prec;
    void* st;
    void* part = nullptr;        // shared-node partials buffer (nullptr: the descriptor's part)
    int part_cap = 0x7fffffff;   // its capacity in entries
    bool tail = false;           // tail-balanced plans: the gather-first product at any launch size
    const dfea_desc* host = nullptr;   // one solve: the host copy of its descriptor (finish pointers as arguments)
    template <typename T, int NEN, int BS0> int go() {
        constexpr bool SMALL = NEN == 4 && BS0 == BS_QUAD;     // small-chunk quad4 plans (see pcg_iter)
        constexpr int CEX = SMALL ? CHUNK_QUAD : default_ce<NEN>();
        constexpr int BS = quadratic<NEN>() ? bs_groups<T>() : SMALL ? BS_Q4S : BS0;
        const T* vv = static_cast<const T*>(v);
        T* yy = static_cast<T*>(y);
        // linear elements: Gauss points requested one step ahead while the launch fits in the workgroup slots
        // of the (costlier in registers) ahead variant; larger quad4 launches start the element stream before
        // the gather instead
        const bool ahead = NEN <= 4 && !SMALL && (n <= APPLY_AHEAD_ITEMS || tail);
        V2<T>* pp = static_cast<V2<T>*>(part);
        constexpr bool AH = NEN <= 4 && !SMALL, SF = NEN == 4 && !SMALL;
        const dfea_chunk* hc = host ? host->chunks : nullptr;
        int rc = host ? (ahead ? launch(k_apply<NEN, nq_of<NEN>(), BS, T, 2, 1, AH, false, CEX, true>, n, BS, st,
                                        "k_apply", hc, t, it, vv, yy, pp, part_cap)
                               : launch(k_apply<NEN, nq_of<NEN>(), BS, T, 2, 1, false, SF, CEX, true>, n, BS, st,
                                        "k_apply", hc, t, it, vv, yy, pp, part_cap))
                      : (ahead ? launch(k_apply<NEN, nq_of<NEN>(), BS, T, 2, 1, AH, false, CEX>, n, BS, st,
                                        "k_apply", hc, t, it, vv, yy, pp, part_cap)
                               : launch(k_apply<NEN, nq_of<NEN>(), BS, T, 2, 1, false, SF, CEX>, n, BS, st,
                                        "k_apply", hc, t, it, vv, yy, pp, part_cap));
        if (rc) return rc;
        if (host)
            return launch(k_apply_fin_direct<T>, nf, FBS, st, "k_apply_fin", host->sh_off, host->sh_node,
                          (const V2<T>*)(pp ? (const void*)pp : (const void*)host->part), vv, yy, host->n_shared);
        return launch(k_apply_fin<T>, nf, FBS, st, "k_apply_fin", t, fi, vv, yy, (const V2<T>*)pp, part_cap);
    }
    template <int NEN, int BS, bool LIB> int run() {
        return prec == DFEA_FP64 ? go<double, NEN, BS>() : go<float, NEN, BS>();
    }
};

struct Launch_pcg_iterate {
    const dfea_desc* t; const int32_t* it; int n; int k; int prec; void* st;
    bool wide = false;
    template <int NEN, int BS, bool LIB> int run() {
        if constexpr (quadratic<NEN>()) {
            if (wide) return prec == DFEA_FP64 ? pcg_iter<NEN, BS, double, true>(t, it, n, k, st)
                                               : pcg_iter<NEN, BS, float, true>(t, it, n, k, st);
        }
        return prec == DFEA_FP64 ? pcg_iter<NEN, BS, double>(t, it, n, k, st)
                                 : pcg_iter<NEN, BS, float>(t, it, n, k, st);
    }
};

struct Launch_adjoint {
    const dfea_desc* t; const int32_t* it; int n; const double* u; const double* lam; double* g; void* st;
    template <int NEN, int BS, bool LIB> int run() {
        return launch(k_adjoint<NEN, nq_of<NEN>(), BS, LIB>, n, BS, st, "k_adjoint", t, it, u, lam, g);
    }
};

}  // namespace

// ======================================================================================================
// scratch of the no-allocation product: the shared-node partials of dfea_apply_noalloc live in this array,
// one copy per device (each GPU the library runs on holds it until the process ends).  Host code never uses
// the array's name as an address: the device address of the current device's copy comes from
// hipGetSymbolAddress, once per device, and is kept in g_scratch[].
// ======================================================================================================
static __device__ __align__(256) unsigned char scratch[DFEA_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch[MAX_DEVICES] = {};

// device address of the current device's scratch (looked up on the first call on that device)
int scratch_address(void** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(-1000 - (int)e, "dfea_apply_noalloc: hipGetDevice failed");
    if (dev < 0 || dev >= MAX_DEVICES) return fail(DFEA_EINVAL, "dfea_apply_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(-1000 - (int)e, "dfea_apply_noalloc: hipGetSymbolAddress(scratch) failed");
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return 0;
}

}  // namespace

// ======================================================================================================
// C ABI
// ======================================================================================================
extern "C" {

int dfea_version(void) { return 100; }
const char* dfea_last_error(void) { return g_err; }
int dfea_desc_size(void) { return (int)sizeof(dfea_desc); }

int dfea_desc_offset(int field) {
    static const int off[] = {
        (int)offsetof(dfea_desc, chunks), (int)offsetof(dfea_desc, lnode), (int)offsetof(dfea_desc, lelem),
        (int)offsetof(dfea_desc, sh_node), (int)offsetof(dfea_desc, sh_off), (int)offsetof(dfea_desc, pslot), (int)offsetof(dfea_desc, hel), (int)offsetof(dfea_desc, hlelem),
        (int)offsetof(dfea_desc, prange), (int)offsetof(dfea_desc, inv_off), (int)offsetof(dfea_desc, inv_slot),
        (int)offsetof(dfea_desc, conn), (int)offsetof(dfea_desc, perm), (int)offsetof(dfea_desc, nodes),
        (int)offsetof(dfea_desc, dNdxi), (int)offsetof(dfea_desc, gw), (int)offsetof(dfea_desc, nfix),
        (int)offsetof(dfea_desc, dN64), (int)offsetof(dfea_desc, dN32), (int)offsetof(dfea_desc, w),
        (int)offsetof(dfea_desc, theta), (int)offsetof(dfea_desc, C64), (int)offsetof(dfea_desc, C32),
        (int)offsetof(dfea_desc, diag64), (int)offsetof(dfea_desc, diag32), (int)offsetof(dfea_desc, f),
        (int)offsetof(dfea_desc, part), (int)offsetof(dfea_desc, red), (int)offsetof(dfea_desc, sc),
        (int)offsetof(dfea_desc, vec), (int)offsetof(dfea_desc, mat_p),
        (int)offsetof(dfea_desc, n_chunks), (int)offsetof(dfea_desc, n_lnodes), (int)offsetof(dfea_desc, n_shared),
        (int)offsetof(dfea_desc, n_part), (int)offsetof(dfea_desc, n_el), (int)offsetof(dfea_desc, n_nodes),
        (int)offsetof(dfea_desc, nen), (int)offsetof(dfea_desc, nq), (int)offsetof(dfea_desc, precision),
        (int)offsetof(dfea_desc, n_finblk), (int)offsetof(dfea_desc, n_vblk), (int)offsetof(dfea_desc, n_halo),
        (int)offsetof(dfea_desc, n_cols), (int)offsetof(dfea_desc, precond), (int)offsetof(dfea_desc, material),
        (int)offsetof(dfea_desc, mat_scal), (int)offsetof(dfea_desc, lstride)};
    if (field < 0 || field >= (int)(sizeof off / sizeof off[0])) return -1;
    return off[field];
}

int dfea_red_size(int n_chunks, int n_finblk, int n_vblk) { return 6 * n_chunks + 6 + n_finblk + 3 * n_vblk; }
int dfea_vblk(int n_dof) { return (n_dof + VBLK - 1) / VBLK; }
int dfea_finblk(int n_shared) { return (n_shared + FBS - 1) / FBS; }
int dfea_chunk_elems(int nen) {
    switch (kind_of(nen)) {
        case Kind::T3: case Kind::Q4: case Kind::Q4H: return CHUNK_LIN;
        case Kind::Q4S: case Kind::T6: case Kind::Q8: case Kind::Q9: return CHUNK_QUAD;
        default: return -1;
    }
}
int dfea_chunk_nodes(void) { return MAXLN; }
int dfea_chunk_threads(int nen) {
    switch (kind_of(nen)) {
        case Kind::T3: case Kind::Q4: return BS_LIN;
        case Kind::Q4H: return BS_Q4H;
        case Kind::Q4S: case Kind::T6: case Kind::Q8: case Kind::Q9: return BS_QUAD;
        default: return -1;
    }
}

int dfea_geometry(const dfea_desc* table, const int32_t* items, int n_items, int nen, void* stream) {
    if (!table) return fail(DFEA_EINVAL, "dfea_geometry: bad argument");
    return dispatch(nen, Launch_geometry{table, items, n_items, stream});
}

int dfea_moduli(const dfea_desc* table, const int32_t* items, int n_items, const int32_t* fin_items,
                int n_fin_items, const double* u, int nen, int precision, void* stream) {
    if (!table || (precision != DFEA_FP64 && precision != DFEA_FP32))
        return fail(DFEA_EINVAL, "dfea_moduli: bad argument");
    return dispatch(nen, Launch_moduli{table, items, n_items, fin_items, n_fin_items, u, precision, stream});
}

int dfea_residual(const dfea_desc* table, const int32_t* items, int n_items, const int32_t* fin_items,
                  int n_fin_items, const double* u, double* R, double lf, int nen, void* stream) {
    if (!table) return fail(DFEA_EINVAL, "dfea_residual: bad argument");
    return dispatch(nen, Launch_residual{table, items, n_items, fin_items, n_fin_items, u, R, lf, stream});
}

int dfea_reactions(const dfea_desc* table, const int32_t* items, int n_items, const int32_t* solve_items,
                   int n_solves, const double* u, double* out, int nen, void* stream) {
    if (!table || !out || n_solves < 1) return fail(DFEA_EINVAL, "dfea_reactions: bad argument");
    return dispatch(nen, Launch_reactions{table, items, n_items, solve_items, n_solves, u, out, stream});
}

int dfea_fields(const dfea_desc* table, const int32_t* items, int n_items, const int32_t* fin_items,
                int n_fin_items, const double* u, double* eout, double* nout, double* work, int nen, void* stream) {
    if (!table || !eout || !nout || (n_fin_items > 0 && !work)) return fail(DFEA_EINVAL, "dfea_fields: bad argument");
    return dispatch(nen, Launch_fields{table, items, n_items, fin_items, n_fin_items, u, eout, nout, work, stream});
}

int dfea_apply(const dfea_desc* table, const int32_t* items, int n_items, const int32_t* fin_items,
               int n_fin_items, const void* v, void* y, int nen, int precision, void* stream) {
    if (!table || (precision != DFEA_FP64 && precision != DFEA_FP32))
        return fail(DFEA_EINVAL, "dfea_apply: bad argument");
    Launch_apply la{table, items, n_items, fin_items, n_fin_items, v, y, precision, stream};
    la.tail = (nen & DFEA_TAIL) != 0;
    return dispatch(nen, la);
}

int dfea_apply_host(const dfea_desc* host_desc, const dfea_desc* table, int n_chunks, int n_fin_blocks, const void* v,
                    void* y, int nen, int precision, void* stream) {
    if (!host_desc || !table || !v || !y || n_chunks < 0 || (precision != DFEA_FP64 && precision != DFEA_FP32) ||
        n_fin_blocks != (host_desc->n_shared + FBS - 1) / FBS || n_chunks != host_desc->n_chunks)
        return fail(DFEA_EINVAL, "dfea_apply_host: bad argument");
    Launch_apply la{table, nullptr, n_chunks, nullptr, n_fin_blocks, v, y, precision, stream};
    la.tail = (nen & DFEA_TAIL) != 0;
    la.host = host_desc;
    return dispatch(nen, la);
}

int dfea_apply_noalloc(const dfea_desc* table, int n_chunks, int n_fin_blocks, int n_part, const void* v, void* y,
                       int nen, int precision, void* stream) {
    if (!table || !v || !y || n_chunks < 0 || n_fin_blocks < 0 || n_part < 0 ||
        (precision != DFEA_FP64 && precision != DFEA_FP32))
        return fail(DFEA_EINVAL, "dfea_apply_noalloc: bad argument");
    const size_t entry = precision == DFEA_FP64 ? 2 * sizeof(double) : 2 * sizeof(float);
    if ((size_t)n_part * entry > (size_t)DFEA_NOALLOC_SCRATCH_BYTES) {
        snprintf(g_err, sizeof g_err,
                 "dfea_apply_noalloc: %d shared-node partials need %zu bytes, more than the %d-byte scratch "
                 "(use dfea_apply with the descriptor's partials buffer)",
                 n_part, (size_t)n_part * entry, DFEA_NOALLOC_SCRATCH_BYTES);
        return DFEA_EINVAL;
    }
    void* part = nullptr;
    const int rc = scratch_address(&part);
    if (rc) return rc;
    Launch_apply la{table, nullptr, n_chunks, nullptr, n_fin_blocks, v, y, precision, stream};
    la.part = part;
    la.part_cap = n_part;
    la.tail = (nen & DFEA_TAIL) != 0;
    return dispatch(nen, la);
}

int dfea_dot(const dfea_desc* table, const int32_t* items, int n_items, const int32_t* solve_items, int n_solves,
             const double* x, const double* y, void* stream) {
    if (!table) return fail(DFEA_EINVAL, "dfea_dot: bad argument");
    int rc = launch(k_dot_part, n_items, VBS, stream, "k_dot_part", table, items, x, y);
    if (rc) return rc;
    return launch(k_dot_final, n_solves, VBS, stream, "k_dot_final", table, solve_items);
}

int dfea_axpy(const dfea_desc* table, const int32_t* items, int n_items, int mode, const double* u, const void* dx,
              const double* uD, const double* sarg, double* out, int precision, void* stream) {
    if (!table || (mode != 0 && mode != 1 && mode != 2) || (mode != 2 && !sarg))
        return fail(DFEA_EINVAL, "dfea_axpy: bad argument");
    if (precision == DFEA_FP64)
        return launch(k_axpy<double>, n_items, VBS, stream, "k_axpy", table, items, mode, u,
                      static_cast<const double*>(dx), uD, sarg, out);
    return launch(k_axpy<float>, n_items, VBS, stream, "k_axpy", table, items, mode, u,
                  static_cast<const float*>(dx), uD, sarg, out);
}

int dfea_pcg_init(const dfea_desc* table, const int32_t* vitems, int n_vitems, const int32_t* solve_items,
                  int n_solves, int neg_res, int precision, void* stream) {
    if (!table) return fail(DFEA_EINVAL, "dfea_pcg_init: bad argument");
    int rc = precision == DFEA_FP64
                 ? launch(k_pcg_init<double>, n_vitems, VBS, stream, "k_pcg_init", table, vitems, neg_res)
                 : launch(k_pcg_init<float>, n_vitems, VBS, stream, "k_pcg_init", table, vitems, neg_res);
    if (rc) return rc;
    return launch(k_pcg_init_final, n_solves, VBS, stream, "k_pcg_init_final", table, solve_items, 0, 0.0, 0.0);
}

int dfea_pcg_start(const dfea_desc* table, int n_vitems, double rtol, int max_iter, int neg_res, int precision,
                   void* stream) {
    if (!table || n_vitems <= 0 || max_iter < 0) return fail(DFEA_EINVAL, "dfea_pcg_start: bad argument");
    const int* none = nullptr;
    int rc = precision == DFEA_FP64
                 ? launch(k_pcg_init<double>, n_vitems, VBS, stream, "k_pcg_init", table, none, neg_res)
                 : launch(k_pcg_init<float>, n_vitems, VBS, stream, "k_pcg_init", table, none, neg_res);
    if (rc) return rc;
    return launch(k_pcg_init_final, 1, VBS, stream, "k_pcg_init_final", table, none, 1, rtol, (double)max_iter);
}

int dfea_pcg_iterate(const dfea_desc* table, const int32_t* items, int n_items, int n_iter, int nen, int precision,
                     void* stream) {
    if (!table || n_iter < 0 || (precision != DFEA_FP64 && precision != DFEA_FP32))
        return fail(DFEA_EINVAL, "dfea_pcg_iterate: bad argument");
    Launch_pcg_iterate lp{table, items, n_items, n_iter, precision, stream};
    lp.wide = (nen & DFEA_WIDE) != 0;
    return dispatch(nen, lp);
}

int dfea_pcg_graph_create(const dfea_desc* table, const int32_t* items, int n_items, int n_iter, int nen,
                          int precision, void** graph) {
    if (!graph || n_iter <= 0) return fail(DFEA_EINVAL, "dfea_pcg_graph_create: bad argument");
    *graph = nullptr;
    hipStream_t s = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(-1000 - (int)e, "dfea_pcg_graph_create: hipStreamCreate failed");
    e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(s);
        return fail(-1000 - (int)e, "dfea_pcg_graph_create: hipStreamBeginCapture failed");
    }
    int rc = dfea_pcg_iterate(table, items, n_items, n_iter, nen, precision, s);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s, &g);
    (void)hipStreamDestroy(s);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess || !g) return fail(-1000 - (int)e, "dfea_pcg_graph_create: hipStreamEndCapture failed");
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(-1000 - (int)e, "dfea_pcg_graph_create: hipGraphInstantiate failed");
    *graph = (void*)x;
    return 0;
}

int dfea_graph_launch(void* graph, void* stream) {
    if (!graph) return fail(DFEA_EINVAL, "dfea_graph_launch: null graph");
    hipError_t e = hipGraphLaunch((hipGraphExec_t)graph, S(stream));
    if (e != hipSuccess) return fail(-1000 - (int)e, "dfea_graph_launch: hipGraphLaunch failed");
    return 0;
}

int dfea_graph_destroy(void* graph) {
    if (!graph) return 0;
    hipError_t e = hipGraphExecDestroy((hipGraphExec_t)graph);
    if (e != hipSuccess) return fail(-1000 - (int)e, "dfea_graph_destroy failed");
    return 0;
}

int dfea_adjoint_grad(const dfea_desc* table, const int32_t* items, int n_items, const double* u, const double* lam,
                      double* g, int nen, void* stream) {
    if (!table || !u || !lam || !g) return fail(DFEA_EINVAL, "dfea_adjoint_grad: bad argument");
    return dispatch(nen, Launch_adjoint{table, items, n_items, u, lam, g, stream});
}

}  // extern "C"
