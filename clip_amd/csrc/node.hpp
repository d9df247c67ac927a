// node.hpp — per-node arithmetic of the coupled step in moment space, shared
// by the FP32 / mixed step kernels, the reverse-mode (adjoint) kernels and a
// host build used by the CPU tests.
//
// The WMRT collision of kernelCollisionMRTf (NsAllen.cu:589-763) relaxes only
// the stress moments (D2Q9 rows {7,8}, D3Q19 rows {4..8}), and those rows
// span the traceless second-order moments.  With the lattice's fourth-order
// isotropy (sum w e_a e_b e_c e_d = (dd + dd + dd)/9) this gives, exactly:
//   Pi(gneq)            = Pi(f_post) - p/3 I - u u^T                (= P)
//   Pi(filter(gneq))    = P + (s - 1) dev(P)
//   sum_k M^-1[q][k] m_k = 9/2 w_q (e_q e_q : dev(P))
//   f*_q = p w_q + Gamma_q(u) + 3/2 w_q (e_q . F) / rho - 9/2 (s-1) w_q (e_q e_q : dev(P))
// with the forcing term's stress moments vanishing (odd in e).  The viscous
// force of calculateVF (NsAllen.cu:102-132) is -tau drho Pi(filter(gneq)) grad(phi).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "clip_amd.h"
#include "lattice.hpp"

#define CLIP_HD __host__ __device__ __forceinline__

namespace clipamd {

// Quotient x / d from r = 1/d (r correctly rounded): q0 = x r, residual
// x - q0 d exact by FMA, q = q0 + residual r (Markstein, IBM J. Res. Dev.
// 34(1), 1990).  Equals the IEEE quotient for the operands of this path
// (checked against plain division on random operands by the CPU tests), so
// the FP64 results stay identical to the oracle's while each reuse of a
// divisor costs 3 operations.
template <class T>
CLIP_HD T div_rn(T x, T d, T r) {
    const T q0 = x * r;
    const T rem = fma(-q0, d, x);
    return fma(rem, r, q0);
}

template <class CT>
struct Phys {
    CT rho_h, rho_l, tau_h, tau_l, beta, kappa, mobility, W;
    CT grav[3];
    int case_type;
};

template <class T>
CLIP_HD T tau_at(const Phys<T>& ph, T phi) {
    if (phi < T(0)) return ph.tau_l;
    if (phi > T(1)) return ph.tau_h;
    return ph.tau_l + phi * (ph.tau_h - ph.tau_l);
}

// d(body force)/d(rho) per unit gravity; the force is base(rho) * g
// (RTI: rho, bubble: rho - rho_h, drop: rho - rho_l, others: 0).
template <class T>
CLIP_HD T body_base(const Phys<T>& ph, T rho, T& dbase) {
    switch (ph.case_type) {
        case CLIP_CASE_RTI: dbase = T(1); return rho;
        case CLIP_CASE_BUBBLE: dbase = T(1); return rho - ph.rho_h;
        case CLIP_CASE_DROP: dbase = T(1); return rho - ph.rho_l;
        default: dbase = T(0); return T(0);
    }
}

// Sym(3x3) helpers over the first D axes.
template <int D, class T>
CLIP_HD T trace(const T a[3][3]) {
    T t = T(0);
#pragma unroll
    for (int i = 0; i < D; ++i) t += a[i][i];
    return t;
}
template <int D, class T>
CLIP_HD void deviator(const T a[3][3], T o[3][3]) {
    const T t = trace<D>(a) * (T(1) / T(D));
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) o[i][j] = (i < D && j < D) ? a[i][j] - (i == j ? t : T(0)) : T(0);
}
template <int D, class T>
CLIP_HD T ddot(const T a[3][3], const T b[3][3]) {
    T t = T(0);
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) t += a[i][j] * b[i][j];
    return t;
}

// e_q . v and e_q e_q : A (A symmetric) with compile-time lattice vectors.
template <class L, class T>
CLIP_HD T edot(int q, const T v[3]) {
    T t = T(0);
    if (L::ex(q)) t += T(L::ex(q)) * v[0];
    if (L::ey(q)) t += T(L::ey(q)) * v[1];
    if (L::D == 3 && L::ez(q)) t += T(L::ez(q)) * v[2];
    return t;
}
template <class L, class T>
CLIP_HD T eeA(int q, const T A[3][3]) {
    T t = T(0);
#pragma unroll
    for (int a = 0; a < L::D; ++a) {
        const int ea = L::e(q, a);
        if (!ea) continue;
        t += A[a][a];
#pragma unroll
        for (int b = a + 1; b < L::D; ++b) {
            const int eb = L::e(q, b);
            if (eb) t += T(2 * ea * eb) * A[a][b];
        }
    }
    return t;
}

// Gamma_q(u) = w_q (3 e.u + 9/2 (e.u)^2 - 3/2 u.u)  (Equilibrium_new, NsAllen.cu:77-94)
template <class L, class T>
CLIP_HD T gamma_of(int q, T eu, T uu) {
    return T(L::w(q)) * (T(3) * eu + T(4.5) * eu * eu - T(1.5) * uu);
}

// First and second moments of the pulled populations.
template <class L, class T>
CLIP_HD void pop_moments(const T fp[L::Q], T mom[3], T Pf[3][3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        mom[a] = T(0);
#pragma unroll
        for (int b = 0; b < 3; ++b) Pf[a][b] = T(0);
    }
#pragma unroll
    for (int q = 0; q < L::Q; ++q) {
#pragma unroll
        for (int a = 0; a < L::D; ++a) {
            const int ea = L::e(q, a);
            if (!ea) continue;
            mom[a] += T(ea) * fp[q];
            Pf[a][a] += fp[q];
#pragma unroll
            for (int b = a + 1; b < L::D; ++b) {
                const int eb = L::e(q, b);
                if (eb) Pf[a][b] += T(ea * eb) * fp[q];
            }
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < a; ++b) Pf[a][b] = Pf[b][a];
}

// P = Pf - p/3 I - v v^T, its deviator, and Pi' = P + (s-1) dev(P).
template <class L, class T>
CLIP_HD void stress_of(const T Pf[3][3], T p, const T v[3], T sm1, T P[3][3], T dP[3][3], T Pp[3][3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
            P[a][b] = (a < L::D && b < L::D) ? Pf[a][b] - (a == b ? p * T(1.0 / 3.0) : T(0)) - v[a] * v[b] : T(0);
    deviator<L::D>(P, dP);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) Pp[a][b] = P[a][b] + sm1 * dP[a][b];
}

// Node quantities that do not depend on the populations.
template <class T>
struct NodeScalars {
    T drho, rho, rrho, tau, s, sm1, mu, base, dbase, fg[3], wc, A;
};
template <class L, class T>
CLIP_HD NodeScalars<T> node_scalars(const Phys<T>& ph, T c0, T lap) {
    NodeScalars<T> z;
    z.drho = ph.rho_h - ph.rho_l;
    z.rho = ph.rho_l + c0 * z.drho;
    z.rrho = T(1) / z.rho;
    z.tau = tau_at(ph, c0);
    z.s = T(1) / (z.tau + T(0.5));
    z.sm1 = z.s - T(1);
    z.mu = T(4) * ph.beta * c0 * (c0 - T(1)) * (c0 - T(0.5)) - ph.kappa * lap;
    z.base = body_base(ph, z.rho, z.dbase);
#pragma unroll
    for (int d = 0; d < 3; ++d) z.fg[d] = (d < L::D) ? z.base * ph.grav[d] : T(0);
    z.wc = T(1) / (T(0.5) + T(3) * ph.mobility);
    const T h = c0 - T(0.5);
    z.A = T(3) * ph.mobility * (T(1) - T(4) * h * h) / ph.W;
    return z;
}

// F = (mu - p drho/3) grad(phi) - tau drho Pi' grad(phi) + f_g
template <class L, class T>
CLIP_HD void force_of(const NodeScalars<T>& z, T p, const T dc[3], const T Pp[3][3], T F[3]) {
    const T cF = z.mu - p * (z.drho * T(1.0 / 3.0));
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        T acc = T(0);
#pragma unroll
        for (int b = 0; b < L::D; ++b) acc += Pp[a][b] * dc[b];
        F[a] = (a < L::D) ? cF * dc[a] - z.tau * z.drho * acc + z.fg[a] : T(0);
    }
}

// |dc|.  In FP32, dc . dc underflows for |dc| below ~1e-19 (far-field tails
// of the tanh profile), which would leave n = dc / (|dc| + 1e-32) far from
// unit length; the magnitude is then taken from dc scaled by its largest
// component.
template <class L, class T>
CLIP_HD T magnitude(const T dc[3]) {
    if constexpr (sizeof(T) == 4) {
        T m = fabs(dc[0]) > fabs(dc[1]) ? fabs(dc[0]) : fabs(dc[1]);
        if (L::D == 3 && fabs(dc[2]) > m) m = fabs(dc[2]);
        if (!(m > T(0))) return T(0);
        if (m < T(1e-37)) return m;  // 1/m would overflow; within sqrt(D) of |dc|
        const T r = T(1) / m;
        T sq = (dc[0] * r) * (dc[0] * r) + (dc[1] * r) * (dc[1] * r);
        if (L::D == 3) sq += (dc[2] * r) * (dc[2] * r);
        return m * sqrt(sq);
    } else {
        T sq = dc[0] * dc[0] + dc[1] * dc[1];
        if (L::D == 3) sq += dc[2] * dc[2];
        return sqrt(sq);
    }
}

template <class L, class T>
CLIP_HD void normal_of(const T dc[3], T nr[3], T& rt, T& mag) {
    mag = magnitude<L>(dc);
    rt = T(1) / (mag + T(1e-32));
#pragma unroll
    for (int d = 0; d < 3; ++d) nr[d] = (d < L::D) ? dc[d] * rt : T(0);
}

// macroscopic u of kernelMacroscopicf from the populations' moments and u_old.
template <class L, class T>
CLIP_HD void velocity_of(const NodeScalars<T>& z, T p, const T mom[3], const T Pf[3][3], const T dc[3],
                         const T uo[3], T u[3]) {
    T P[3][3], dP[3][3], Pp[3][3], F[3];
    stress_of<L>(Pf, p, uo, z.sm1, P, dP, Pp);
    force_of<L>(z, p, dc, Pp, F);
#pragma unroll
    for (int d = 0; d < 3; ++d) u[d] = (d < L::D) ? mom[d] + T(0.5) * F[d] * z.rrho : T(0);
}

// Both collisions at one node: g* (kernelCollisionMRTg) and f* (moment form of
// kernelCollisionMRTf) from the pulled g, the populations' moments, u and p.
// Emits each population through the callbacks emit_g(q, v) / emit_f(q, v).
template <class L, class T, class EG, class EF>
CLIP_HD void collide_node(const NodeScalars<T>& z, T p, const T Pf[3][3], T c0, const T dc[3], const T u[3],
                          const T gp[L::Q], EG&& emit_g, EF&& emit_f) {
    T nr[3], rt, mag;
    normal_of<L>(dc, nr, rt, mag);
    T P[3][3], dP[3][3], Pp[3][3], F[3];
    stress_of<L>(Pf, p, u, z.sm1, P, dP, Pp);
    force_of<L>(z, p, dc, Pp, F);
    T uu = u[0] * u[0] + u[1] * u[1];
    if (L::D == 3) uu += u[2] * u[2];
    const T h = T(1.5) * z.rrho;
    const T k = T(4.5) * z.sm1;
    const T omw = T(1) - z.wc;
#pragma unroll
    for (int q = 0; q < L::Q; ++q) {
        const T w = T(L::w(q));
        const T gm = gamma_of<L>(q, edot<L>(q, u), uu);
        const T heq = c0 * (gm + w) + w * (z.A * edot<L>(q, nr));
        emit_g(q, omw * gp[q] + z.wc * heq);
        emit_f(q, p * w + gm + w * (h * edot<L>(q, F) - k * eeA<L>(q, dP)));
    }
}

// ---------------------------------------------------------------- reverse
// Reverse of nr = dc / (|dc| + 1e-32).  Every use of nr is weighted by
// 4 phi (1 - phi) (the interface term A and the initial g), so where |dc| is
// far below the equilibrium slope 4 phi (1 - phi) / W (flat plateaus, the
// centre of a drop) the exact derivative ~ 1/|dc| is noise that grows without
// bound over many steps.  nfloor > 0 scales the derivative by
// |dc| / max(|dc|, nfloor 4 phi (1 - phi) / W), bounding it where the
// profile is degenerate and leaving it exact elsewhere; nfloor = 0 is the
// exact derivative of the step.
template <class L, class T>
CLIP_HD void normal_reverse(const T dc[3], T rt, T mag, T c0, T W, T nfloor, const T lnr[3], T ldc[3]) {
    constexpr int D = L::D;
    T sc = T(1);
    if (nfloor > T(0)) {
        T fl = nfloor * T(4) * c0 * (T(1) - c0) / W;
        fl = fl < T(0) ? -fl : fl;
        sc = (mag > fl) ? T(1) : (fl > T(0) ? mag / fl : T(0));
    }
    // dn_i/ddc_j = rt (delta_ij - n_i m_j), m = dc / |dc|: no rt^2 (which
    // overflows FP32 for |dc| near 0)
    T nl = T(0);
#pragma unroll
    for (int d = 0; d < D; ++d) nl += lnr[d] * dc[d] * rt;
    // 1/|dc| overflows FP32 for subnormal |dc|: drop the (bounded) second
    // term there
    const T im = (mag > T(1e-37)) ? T(1) / mag : T(0);
    const T k = sc * rt;
#pragma unroll
    for (int d = 0; d < D; ++d) ldc[d] += k * (lnr[d] - nl * dc[d] * im);
}

// Adjoint of one node of a full step (velocity from u_old, then both
// collisions): given the adjoints of f*, g* and u, accumulates the adjoints of
// the pulled populations, of phi, grad(phi), lap(phi) at the node, of u_old,
// and of beta, kappa and the mobility.
template <class L, class T>
struct NodeAdj {
    T lfp[L::Q], lgp[L::Q];
    T lc0, ldc[3], llap, luo[3];
    T lbeta, lkappa, lmob;
};

template <int D, class T>
CLIP_HD void sym_add_outer(T M[3][3], T a, const T x[3], const T y[3]) {
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) M[i][j] += a * x[i] * y[j];
}

// Reverse of F = cF dc - tau drho Pp dc + fg (Pp from P = Pf - p/3 I - v v^T):
// adds to lmu, lp, ldc, lfg, ltau, ls, lPf and lv.
template <class L, class T>
CLIP_HD void force_reverse(const NodeScalars<T>& z, T p, const T dc[3], const T v[3], const T Pf[3][3],
                           const T lF[3], T& lmu, T& lp, T ldc[3], T lfg[3], T& ltau, T& ls, T lPf[3][3],
                           T lv[3]) {
    constexpr int D = L::D;
    T P[3][3], dP[3][3], Pp[3][3];
    stress_of<L>(Pf, p, v, z.sm1, P, dP, Pp);
    const T cF = z.mu - p * (z.drho * T(1.0 / 3.0));
    T lcF = T(0);
#pragma unroll
    for (int a = 0; a < D; ++a) {
        lcF += lF[a] * dc[a];
        ldc[a] += cF * lF[a];
        lfg[a] += lF[a];
    }
    lmu += lcF;
    lp -= (z.drho * T(1.0 / 3.0)) * lcF;
    // fv = -tau drho Pp dc
    T lPp[3][3] = {};
    T Ppdc[3] = {T(0), T(0), T(0)};
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) {
            Ppdc[a] += Pp[a][b] * dc[b];
            lPp[a][b] = -z.tau * z.drho * lF[a] * dc[b];
            ldc[b] += -z.tau * z.drho * Pp[a][b] * lF[a];
        }
    T t = T(0);
#pragma unroll
    for (int a = 0; a < D; ++a) t += lF[a] * Ppdc[a];
    ltau += -z.drho * t;
    // Pp = P + sm1 dev(P)
    T dl[3][3];
    deviator<D>(lPp, dl);
    ls += ddot<D>(lPp, dP);
    T lP[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) lP[a][b] = lPp[a][b] + z.sm1 * dl[a][b];
    // P = Pf - p/3 I - v v^T
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) {
            lPf[a][b] += lP[a][b];
            lv[a] -= (lP[a][b] + lP[b][a]) * v[b];
        }
    lp -= trace<D>(lP) * T(1.0 / 3.0);
}

template <class L, class T>
CLIP_HD void node_reverse(const Phys<T>& ph, const T fp[L::Q], const T gp[L::Q], T c0, const T dc[3], T lap,
                          const T uo[3], const T lfs[L::Q], const T lgs[L::Q], const T lu_out[3],
                          NodeAdj<L, T>& out, T nfloor = T(0)) {
    constexpr int D = L::D;
    constexpr int Q = L::Q;
    const NodeScalars<T> z = node_scalars<L>(ph, c0, lap);
    T p = T(0);
#pragma unroll
    for (int q = 0; q < Q; ++q) p += fp[q];
    T mom[3], Pf[3][3];
    pop_moments<L>(fp, mom, Pf);
    // forward intermediates
    T Po[3][3], dPo[3][3], Ppo[3][3], Fo[3];
    stress_of<L>(Pf, p, uo, z.sm1, Po, dPo, Ppo);
    force_of<L>(z, p, dc, Ppo, Fo);
    T u[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) u[d] = (d < D) ? mom[d] + T(0.5) * Fo[d] * z.rrho : T(0);
    T nr[3], rt, mag;
    normal_of<L>(dc, nr, rt, mag);
    T Pn[3][3], dPn[3][3], Ppn[3][3], Fn[3];
    stress_of<L>(Pf, p, u, z.sm1, Pn, dPn, Ppn);
    force_of<L>(z, p, dc, Ppn, Fn);
    T uu = u[0] * u[0] + u[1] * u[1];
    if (D == 3) uu += u[2] * u[2];

    T lu[3] = {lu_out[0], lu_out[1], D == 3 ? lu_out[2] : T(0)};
    T lp = T(0), lmom[3] = {T(0), T(0), T(0)}, lPf[3][3] = {};
    T lrho = T(0), ltau = T(0), ls = T(0), lmu = T(0), lfg[3] = {T(0), T(0), T(0)};
    T lnr[3] = {T(0), T(0), T(0)}, lA = T(0), lwc = T(0);
    T lc0 = T(0), ldc[3] = {T(0), T(0), T(0)}, luo[3] = {T(0), T(0), T(0)};
    const T h = T(1.5) * z.rrho;

    // collisions: f*_q = p w + G_q(u) + w h (e.F) - 9/2 sm1 w (ee:dPn),
    //             g*_q = (1-wc) gp + wc (c0 (G_q + w) + w A (e.nr))
    T lFn[3] = {T(0), T(0), T(0)}, Lam[3][3] = {};
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const T w = T(L::w(q));
        const T eu = edot<L>(q, u);
        const T gm = gamma_of<L>(q, eu, uu);
        const T a = lfs[q];
        const T b = lgs[q];
        const T hh = z.wc * b;
        lp += a * w;
        const T cg = a + c0 * hh;  // weight of dGamma/du
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const T ed = T(L::e(q, d));
            lu[d] += cg * w * (T(3) * ed + T(9) * eu * ed - T(3) * u[d]);
            lFn[d] += a * w * h * ed;
            lnr[d] += z.A * hh * w * ed;
        }
        lrho -= a * w * h * edot<L>(q, Fn) * z.rrho;
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j < D; ++j) Lam[i][j] += a * w * T(L::e(q, i) * L::e(q, j));
        const T en = edot<L>(q, nr);
        const T heq = c0 * (gm + w) + w * z.A * en;
        out.lgp[q] = (T(1) - z.wc) * b;
        lwc += b * (heq - gp[q]);
        lc0 += hh * (gm + w);
        lA += hh * w * en;
    }
    // - 9/2 sm1 w (ee : dev(Pn))
    ls += T(-4.5) * ddot<D>(Lam, dPn);
    T lPn_dev[3][3], lPn[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) lPn_dev[i][j] = T(-4.5) * z.sm1 * Lam[i][j];
    deviator<D>(lPn_dev, lPn);
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            lPf[i][j] += lPn[i][j];
            lu[i] -= (lPn[i][j] + lPn[j][i]) * u[j];
        }
    lp -= trace<D>(lPn) * T(1.0 / 3.0);
    force_reverse<L>(z, p, dc, u, Pf, lFn, lmu, lp, ldc, lfg, ltau, ls, lPf, lu);
    // A, wc -> mobility, c0
    {
        const T hc = c0 - T(0.5);
        out.lmob += lA * T(3) * (T(1) - T(4) * hc * hc) / ph.W;
        lc0 += lA * T(3) * ph.mobility * (T(-8) * hc) / ph.W;
        out.lmob += lwc * (T(-3) * z.wc * z.wc);
    }
    // nr = dc / (|dc| + 1e-32)
    normal_reverse<L>(dc, rt, mag, c0, ph.W, nfloor, lnr, ldc);
    // u = mom + F_o / (2 rho)
    T lFo[3] = {T(0), T(0), T(0)};
    {
        T t = T(0);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            lmom[d] += lu[d];
            lFo[d] = T(0.5) * lu[d] * z.rrho;
            t += lu[d] * Fo[d];
        }
        lrho -= T(0.5) * t * z.rrho * z.rrho;
    }
    force_reverse<L>(z, p, dc, uo, Pf, lFo, lmu, lp, ldc, lfg, ltau, ls, lPf, luo);
    // body force, s, tau, mu, rho
    {
        T t = T(0);
#pragma unroll
        for (int d = 0; d < D; ++d) t += lfg[d] * ph.grav[d];
        lrho += t * z.dbase;
    }
    ltau -= ls * z.s * z.s;
    if (c0 >= T(0) && c0 <= T(1)) lc0 += ltau * (ph.tau_h - ph.tau_l);
    lc0 += lmu * T(4) * ph.beta * (T(3) * c0 * c0 - T(3) * c0 + T(0.5));
    out.llap += -ph.kappa * lmu;
    out.lbeta += lmu * T(4) * c0 * (c0 - T(1)) * (c0 - T(0.5));
    out.lkappa += -lmu * lap;
    lc0 += lrho * z.drho;
    // p, mom, Pf -> f_post
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        T t = lp;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            const int ea = L::e(q, a);
            if (!ea) continue;
            t += T(ea) * lmom[a];
#pragma unroll
            for (int b = 0; b < D; ++b) {
                const int eb = L::e(q, b);
                if (eb) t += T(ea * eb) * lPf[a][b];
            }
        }
        out.lfp[q] = t;
    }
    out.lc0 += lc0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        out.ldc[d] += (d < D) ? ldc[d] : T(0);
        out.luo[d] += (d < D) ? luo[d] : T(0);
    }
}

// Adjoint of the initial state and first collision at one node
// (KernelInitializeDistributions, NsAllen.cu:137-202, then the collision with
// u = 0 and the initial p): f = p w, g = phi w - w 2 phi (1 - phi) / W (e.n),
// p = k phi sigma / (L/2) / (rho/3) (k = +1 drop, -1 bubble, 0 otherwise).
// lpf / lpg (optional): adjoints of the initial f / g themselves (they are
// also the first outflow planes on free-convect boundary nodes).
template <class L, class T>
CLIP_HD void init_reverse(const Phys<T>& ph, T sigma, T half_length, T c0, const T dc[3], T lap, const T lfs[L::Q],
                          const T lgs[L::Q], T& lc0, T ldc[3], T& llap, T& lbeta, T& lkappa, T& lmob, T& lsigma,
                          T nfloor = T(0), const T* lpf = nullptr, const T* lpg = nullptr) {
    constexpr int D = L::D;
    constexpr int Q = L::Q;
    const NodeScalars<T> z = node_scalars<L>(ph, c0, lap);
    const T k = ph.case_type == CLIP_CASE_DROP ? T(1) : (ph.case_type == CLIP_CASE_BUBBLE ? T(-1) : T(0));
    const T p = k * c0 * sigma / half_length / (z.rho / T(3));
    T nr[3], rt, mag;
    normal_of<L>(dc, nr, rt, mag);
    const T cF = z.mu - p * (z.drho * T(1.0 / 3.0));
    T F[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) F[d] = (d < D) ? cF * dc[d] + z.fg[d] : T(0);
    const T h = T(1.5) * z.rrho;
    const T gi = T(2) * c0 * (T(1) - c0) / ph.W;  // g_init = w (phi - gi (e.n))
    T lp = T(0), lF[3] = {T(0), T(0), T(0)}, lrho = T(0), lnr[3] = {T(0), T(0), T(0)};
    T lA = T(0), lwc = T(0), lgi = T(0), lc = T(0);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const T w = T(L::w(q));
        const T a = lfs[q];
        const T b = lgs[q];
        const T en = edot<L>(q, nr);
        // f*_q = p w + w h (e.F); f_q = p w
        lp += a * w;
        if (lpf) lp += lpf[q] * w;
#pragma unroll
        for (int d = 0; d < D; ++d) lF[d] += a * w * h * T(L::e(q, d));
        lrho -= a * w * h * edot<L>(q, F) * z.rrho;
        // g*_q = (1-wc) g_init + wc (c0 w + w A en)
        const T gin = w * (c0 - gi * en);
        const T heq = c0 * w + w * z.A * en;
        lwc += b * (heq - gin);
        const T hh = z.wc * b;
        const T lg = (T(1) - z.wc) * b + (lpg ? lpg[q] : T(0));
        lc += hh * w + lg * w;
        lA += hh * w * en;
        lgi -= lg * w * en;
#pragma unroll
        for (int d = 0; d < D; ++d) lnr[d] += (z.A * hh * w - lg * w * gi) * T(L::e(q, d));
    }
    // F = cF dc + fg
    T lmu = T(0), lcF = T(0);
#pragma unroll
    for (int d = 0; d < D; ++d) {
        lcF += lF[d] * dc[d];
        ldc[d] += cF * lF[d];
        lrho += lF[d] * ph.grav[d] * z.dbase;
    }
    lmu += lcF;
    lp -= (z.drho * T(1.0 / 3.0)) * lcF;
    // gi = 2 c0 (1 - c0) / W
    lc += lgi * T(2) * (T(1) - T(2) * c0) / ph.W;
    // A, wc
    {
        const T hc = c0 - T(0.5);
        lmob += lA * T(3) * (T(1) - T(4) * hc * hc) / ph.W;
        lc += lA * T(3) * ph.mobility * (T(-8) * hc) / ph.W;
        lmob += lwc * (T(-3) * z.wc * z.wc);
    }
    // nr
    normal_reverse<L>(dc, rt, mag, c0, ph.W, nfloor, lnr, ldc);
    // p = 3 k c0 sigma / (half_length rho)
    {
        const T c = T(3) * k / half_length;
        lc += lp * c * sigma * z.rrho;
        lrho -= lp * c * sigma * c0 * z.rrho * z.rrho;
        lsigma += lp * c * c0 * z.rrho;
    }
    lc += lmu * T(4) * ph.beta * (T(3) * c0 * c0 - T(3) * c0 + T(0.5));
    llap += -ph.kappa * lmu;
    lbeta += lmu * T(4) * c0 * (c0 - T(1)) * (c0 - T(0.5));
    lkappa += -lmu * lap;
    lc += lrho * z.drho;
    lc0 += lc;
}

}  // namespace clipamd
