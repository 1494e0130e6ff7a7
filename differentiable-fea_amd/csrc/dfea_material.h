// Constitutive closed forms of the material library in plane strain (F3 = diag(F, 1)), device side.
//
// Every energy of `src/difffea/materials.py:65-162` is written out with its first Piola stress P = dpsi/dF
// (2x2, flat index 2i+j), its moduli d2psi/dF_ij dF_kl (symmetric, 10 unique entries in the order
// (00,01,02,03,11,12,13,22,23,33) of the flat index), its out-of-plane stress P33 = dpsi3/dF33 at F33 = 1 and
// its value.  The isotropic energies (neo-Hookean, Mooney-Rivlin, Demiray, the HGO matrix) are functions of
// s = F:F and l = ln J in plane strain, so P = 2 psi_s F + psi_l G and
//   C_ijkl = 2 psi_s d_ik d_jl + 4 psi_ss F_ij F_kl + 2 psi_sl (F_ij G_kl + G_ij F_kl) + psi_ll G_ij G_kl
//            - psi_l G_il G_kj,                                   G = F^-T;
// Saint Venant-Kirchhoff is written through S = lam tr(E) I + 2 mu E: P = F S and
//   C_ijkl = d_ik S_jl + lam F_ij F_kl + mu (F_il F_kj + (F F^T)_ik d_jl);
// the HGO fibre families add W(E_f), E_f = d I1 + (1 - 3d) |F a_f|^2 - 1, W = k1/(2 k2) (exp(k2 <E_f>^2) - 1).
#pragma once

#include <hip/hip_runtime.h>

#include "dfea.h"

namespace dfea_mat {

// LIB: the kernel instantiation carries the whole library; without it only the neo-Hookean closed forms are
// compiled (fewer registers), and a descriptor of another material gets NaN parameters, so its results are
// NaN rather than silently those of another energy.

// the parameters of one element: nominal values, theta-scaled where the material marks them scalable
struct Mat {
    int id;
    double p[DFEA_MAT_NP];
};

template <bool LIB>
__device__ __forceinline__ Mat element_material(const dfea_desc& D, double th) {
    Mat m;
    m.id = D.material;
    constexpr int np = LIB ? DFEA_MAT_NP : 2;
#pragma unroll
    for (int k = 0; k < np; ++k) m.p[k] = D.mat_p[k] * (((D.mat_scal >> k) & 1) ? th : 1.0);
#pragma unroll
    for (int k = np; k < DFEA_MAT_NP; ++k) m.p[k] = 0.0;
    if (!LIB && m.id != DFEA_MAT_NEO_HOOKEAN) m.p[0] = m.p[1] = __builtin_nan("");
    return m;
}

struct Kin2 {
    double F[4], G[4];
    double J, l, s;
};

__device__ __forceinline__ Kin2 kin2(const double* F) {
    Kin2 k;
#pragma unroll
    for (int i = 0; i < 4; ++i) k.F[i] = F[i];
    k.J = F[0] * F[3] - F[1] * F[2];
    const double iJ = 1.0 / k.J;
    k.G[0] = F[3] * iJ;  k.G[1] = -F[2] * iJ;
    k.G[2] = -F[1] * iJ; k.G[3] = F[0] * iJ;
    k.l = log(k.J);
    k.s = F[0] * F[0] + F[1] * F[1] + F[2] * F[2] + F[3] * F[3];
    return k;
}

// derivatives of an isotropic plane-strain energy psi(s, l) and its out-of-plane stress and value
struct Iso {
    double ps, pl, pss, psl, pll, p33, psi;
};

template <bool LIB>
__device__ __forceinline__ Iso iso(const Mat& m, const Kin2& k) {
    Iso d;
    const double s = k.s, l = k.l;
    if (LIB && m.id == DFEA_MAT_MOONEY_RIVLIN) {
        // psi = c10 (J^-2/3 (s + 1) - 3) + c01 (J^-4/3 (J^2 + s) - 3) + kappa/2 l^2
        const double c10 = m.p[0], c01 = m.p[1], kap = m.p[2];
        const double a = exp(-2.0 / 3.0 * l), b = exp(-4.0 / 3.0 * l), c = exp(2.0 / 3.0 * l);
        d.ps = c10 * a + c01 * b;
        d.pl = -2.0 / 3.0 * c10 * a * (s + 1.0) + c01 * (2.0 / 3.0 * c - 4.0 / 3.0 * b * s) + kap * l;
        d.pss = 0.0;
        d.psl = -2.0 / 3.0 * c10 * a - 4.0 / 3.0 * c01 * b;
        d.pll = 4.0 / 9.0 * c10 * a * (s + 1.0) + c01 * (4.0 / 9.0 * c + 16.0 / 9.0 * b * s) + kap;
        d.p33 = c10 * a * (4.0 - 2.0 * s) / 3.0 + c01 * (2.0 / 3.0 * b * s - 4.0 / 3.0 * c) + kap * l;
        d.psi = c10 * (a * (s + 1.0) - 3.0) + c01 * (c + b * s - 3.0) + 0.5 * kap * l * l;
    } else if (LIB && m.id == DFEA_MAT_DEMIRAY) {
        // psi = mu/(2 gamma) (exp(gamma (s - 2)) - 1) - mu l + lam/2 l^2
        const double mu = m.p[0], gam = m.p[1], lam = m.p[2];
        const double ex = exp(gam * (s - 2.0));
        d.ps = 0.5 * mu * ex;
        d.pl = lam * l - mu;
        d.pss = 0.5 * mu * gam * ex;
        d.psl = 0.0;
        d.pll = lam;
        d.p33 = mu * ex - mu + lam * l;
        d.psi = mu / (2.0 * gam) * (ex - 1.0) - mu * l + 0.5 * lam * l * l;
    } else {
        // neo-Hookean (also the HGO matrix): psi = mu/2 (s - 2 - 2 l) + lam/2 l^2
        const double mu = m.p[0], lam = m.p[1];
        d.ps = 0.5 * mu;
        d.pl = lam * l - mu;
        d.pss = 0.0;
        d.psl = 0.0;
        d.pll = lam;
        d.p33 = lam * l;
        d.psi = 0.5 * mu * (s - 2.0 - 2.0 * l) + 0.5 * lam * l * l;
    }
    return d;
}

// Saint Venant-Kirchhoff: S (symmetric 2x2: S00, S01, S11) and tr E
struct Svk {
    double S00, S01, S11, trE, psi;
};

__device__ __forceinline__ Svk svk(const Mat& m, const double* F) {
    const double mu = m.p[0], lam = m.p[1];
    const double E00 = 0.5 * (F[0] * F[0] + F[2] * F[2] - 1.0);
    const double E01 = 0.5 * (F[0] * F[1] + F[2] * F[3]);
    const double E11 = 0.5 * (F[1] * F[1] + F[3] * F[3] - 1.0);
    Svk v;
    v.trE = E00 + E11;
    v.S00 = lam * v.trE + 2.0 * mu * E00;
    v.S01 = 2.0 * mu * E01;
    v.S11 = lam * v.trE + 2.0 * mu * E11;
    v.psi = 0.5 * lam * v.trE * v.trE + mu * (E00 * E00 + 2.0 * E01 * E01 + E11 * E11);
    return v;
}

// one HGO fibre family with direction (a0, a1) and weight wf: E, W'(E), W''(E), W(E) (zero in compression)
struct Fibre {
    double D[4];      // dE/dF
    double w1, w2, W;
    double a0, a1;
};

__device__ __forceinline__ Fibre fibre(const Mat& m, const Kin2& k, double a0, double a1, double wf) {
    const double k1 = m.p[2], k2 = m.p[3], dd = m.p[5];
    Fibre f;
    f.a0 = a0;
    f.a1 = a1;
    const double fa0 = k.F[0] * a0 + k.F[1] * a1, fa1 = k.F[2] * a0 + k.F[3] * a1;
    const double E = dd * (k.s + 1.0) + (1.0 - 3.0 * dd) * (fa0 * fa0 + fa1 * fa1) - 1.0;
    const double g = 2.0 * (1.0 - 3.0 * dd);
    f.D[0] = 2.0 * dd * k.F[0] + g * fa0 * a0;
    f.D[1] = 2.0 * dd * k.F[1] + g * fa0 * a1;
    f.D[2] = 2.0 * dd * k.F[2] + g * fa1 * a0;
    f.D[3] = 2.0 * dd * k.F[3] + g * fa1 * a1;
    if (E > 0.0) {
        const double ex = exp(k2 * E * E);
        f.w1 = wf * k1 * E * ex;
        f.w2 = wf * k1 * ex * (1.0 + 2.0 * k2 * E * E);
        f.W = wf * k1 / (2.0 * k2) * (ex - 1.0);
    } else {
        f.w1 = f.w2 = f.W = 0.0;
    }
    return f;
}

__device__ __forceinline__ void fibre_dirs(const Mat& m, double& c, double& s) {
    const double beta = m.p[4] * (3.14159265358979323846 / 180.0);
    c = cos(beta);
    s = sin(beta);
}

// P (2x2, flat) at F
template <bool LIB>
__device__ __forceinline__ void piola(const Mat& m, const double* F, double* P) {
    if (LIB && m.id == DFEA_MAT_SAINT_VENANT) {
        const Svk v = svk(m, F);
        P[0] = F[0] * v.S00 + F[1] * v.S01;
        P[1] = F[0] * v.S01 + F[1] * v.S11;
        P[2] = F[2] * v.S00 + F[3] * v.S01;
        P[3] = F[2] * v.S01 + F[3] * v.S11;
        return;
    }
    const Kin2 k = kin2(F);
    const Iso d = iso<LIB>(m, k);
#pragma unroll
    for (int i = 0; i < 4; ++i) P[i] = 2.0 * d.ps * k.F[i] + d.pl * k.G[i];
    if (LIB && m.id == DFEA_MAT_FIBER_REINFORCED) {
        double cb, sb;
        fibre_dirs(m, cb, sb);
        const Fibre f1 = fibre(m, k, cb, sb, 1.0), f2 = fibre(m, k, cb, -sb, m.p[6]);
#pragma unroll
        for (int i = 0; i < 4; ++i) P[i] += f1.w1 * f1.D[i] + f2.w1 * f2.D[i];
    }
}

// the 10 unique moduli times w at F
template <bool LIB>
__device__ __forceinline__ void moduli(const Mat& m, const double* F, double w, double* c) {
    if (LIB && m.id == DFEA_MAT_SAINT_VENANT) {
        const double mu = m.p[0], lam = m.p[1];
        const Svk v = svk(m, F);
        const double S[4] = {v.S00, v.S01, v.S01, v.S11};
        const double B[4] = {F[0] * F[0] + F[1] * F[1], F[0] * F[2] + F[1] * F[3], F[0] * F[2] + F[1] * F[3],
                             F[2] * F[2] + F[3] * F[3]};
        int t = 0;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
#pragma unroll
            for (int b = a; b < 4; ++b) {
                const int i = a >> 1, j = a & 1, kk = b >> 1, l = b & 1;
                double v2 = lam * F[a] * F[b] + mu * F[2 * i + l] * F[2 * kk + j];
                if (i == kk) v2 += S[2 * j + l];
                if (j == l) v2 += mu * B[2 * i + kk];
                c[t++] = w * v2;
            }
        }
        return;
    }
    const Kin2 k = kin2(F);
    const Iso d = iso<LIB>(m, k);
    const bool fib = LIB && m.id == DFEA_MAT_FIBER_REINFORCED;
    Fibre f1, f2;
    if (fib) {
        double cb, sb;
        fibre_dirs(m, cb, sb);
        f1 = fibre(m, k, cb, sb, 1.0);
        f2 = fibre(m, k, cb, -sb, m.p[6]);
    }
    int t = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
#pragma unroll
        for (int b = a; b < 4; ++b) {
            const int i = a >> 1, j = a & 1, kk = b >> 1, l = b & 1;
            double v = 4.0 * d.pss * k.F[a] * k.F[b] + 2.0 * d.psl * (k.F[a] * k.G[b] + k.G[a] * k.F[b]) +
                       d.pll * k.G[a] * k.G[b] - d.pl * k.G[2 * i + l] * k.G[2 * kk + j];
            if (i == kk && j == l) v += 2.0 * d.ps;
            if (fib) {
                v += f1.w2 * f1.D[a] * f1.D[b] + f2.w2 * f2.D[a] * f2.D[b];
                if (i == kk) {
                    const double dd = m.p[5], g = 2.0 * (1.0 - 3.0 * dd);
                    const double aa1 = (j == 0 ? f1.a0 : f1.a1) * (l == 0 ? f1.a0 : f1.a1);
                    const double aa2 = (j == 0 ? f2.a0 : f2.a1) * (l == 0 ? f2.a0 : f2.a1);
                    v += f1.w1 * ((j == l ? 2.0 * dd : 0.0) + g * aa1) + f2.w1 * ((j == l ? 2.0 * dd : 0.0) + g * aa2);
                }
            }
            c[t++] = w * v;
        }
    }
}

// P (2x2), the out-of-plane stress P33 and the energy at F (post-processing)
template <bool LIB>
__device__ __forceinline__ void stress_state(const Mat& m, const double* F, double* P, double& P33, double& psi) {
    if (LIB && m.id == DFEA_MAT_SAINT_VENANT) {
        const Svk v = svk(m, F);
        P[0] = F[0] * v.S00 + F[1] * v.S01;
        P[1] = F[0] * v.S01 + F[1] * v.S11;
        P[2] = F[2] * v.S00 + F[3] * v.S01;
        P[3] = F[2] * v.S01 + F[3] * v.S11;
        P33 = m.p[1] * v.trE;
        psi = v.psi;
        return;
    }
    const Kin2 k = kin2(F);
    const Iso d = iso<LIB>(m, k);
#pragma unroll
    for (int i = 0; i < 4; ++i) P[i] = 2.0 * d.ps * k.F[i] + d.pl * k.G[i];
    P33 = d.p33;
    psi = d.psi;
    if (LIB && m.id == DFEA_MAT_FIBER_REINFORCED) {
        double cb, sb;
        fibre_dirs(m, cb, sb);
        const Fibre f1 = fibre(m, k, cb, sb, 1.0), f2 = fibre(m, k, cb, -sb, m.p[6]);
#pragma unroll
        for (int i = 0; i < 4; ++i) P[i] += f1.w1 * f1.D[i] + f2.w1 * f2.D[i];
        P33 += 2.0 * m.p[5] * (f1.w1 + f2.w1);
        psi += f1.W + f2.W;
    }
}

}  // namespace dfea_mat
