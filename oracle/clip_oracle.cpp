// clip_oracle.cpp — FP64 CPU oracle of CLIP's coupled two-phase LBM step.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, phase by phase and in CLIP's
// array-of-structures layout, what the reference's kernels compute.  It is the
// checker of the HIP implementation (tests/, smoke(), bench.py's CPU baseline)
// and is never linked into or called by the product library.
//
// Each phase runs over whole arrays in the reference's order
// (NSAllen::solve, src/equation/NsAllen.cu:881-892), so that the GPU path — which
// fuses the phases into pull-streaming kernels — is checked against an
// unfused statement of the same semantics.  Where the reference is racy or
// reads uninitialised memory the oracle uses the defined semantics below:
//   * periodic ghost copy: sequential X, then Y, then Z passes over whole
//     planes (full periodic wrap at edges and corners)        Solver.cu:75-141
//   * mirror copy of phi: sequential X, Y, Z passes            Solver.cu:927-978
//   * velocity inlet: sides processed X-, X+, Y-, Y+, Z-, Z+ in turn; each
//     side uses its own `value` (the reference reads the y+ value for every
//     side, Solver.cu:587)                                     Solver.cu:550-921
//   * free-convect and Neumann fix-ups read the post-streaming populations of
//     the inner neighbour as they were before the fix-up pass (Jacobi)
//                                                              Solver.cu:342-542
//   * initial pressure has base 0 and the initial normal is 0 at ghost nodes
//     (the reference reads uninitialised device memory)        NsAllen.cu:137-202
//   * jet case: gravity body force 0 (the reference leaves it unset)
//                                                              NsAllen.cu:460-474
//
// Build: oracle/build.sh (g++ -O2 -fopenmp -shared).
#include "clip_amd.h"

// Planted mistakes for the mutation checks of the pin tests
// (tests/test_oracle_mutants.py builds the oracle with -DCLIP_ORACLE_MUTANT=k
// and expects the pin of the mutated function to fail).  0: none.
#ifndef CLIP_ORACLE_MUTANT
#define CLIP_ORACLE_MUTANT 0
#endif
#define MUTANT(k) (CLIP_ORACLE_MUTANT == (k))

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

// ---------------------------------------------------------------- lattice
struct Lattice {
    int D = 2, Q = 9;
    int e[19][3];
    double w[19];
    int opp[19];
    int nstress = 0;
    int stress[5];
    double M[19][19];
    double Minv[19][19];
    // M^-1 as integer numerators over one denominator per row (the form the
    // reference's reconvert* uses, WMRT.cuh:77-92, 134-184), recovered from
    // the numerical inverse.
    double Minv_num[19][19];
    double Minv_den[19];
    // in-directions of each side (pointing from the side into the domain),
    // axis-normal direction first: WMRT::wallBCMap (WMRT.cuh:214-237)
    int nin = 0;
    int in[6][5];
    int weight = 0;  // number of non-normal in-directions (wallBCMap::weight)
};

// Moment basis polynomials, evaluated at lattice velocity (x, y, z).  They
// define CLIP's convert*Weighted rows (WMRT.cuh:54-126).
double moment2(int r, int x, int y) {
    const double e2 = x * x + y * y;
    switch (r) {
        case 0: return 1;
        case 1: return 3 * e2 - 4;
        case 2: return 4.5 * e2 * e2 - 10.5 * e2 + 4;
        case 3: return x;
        case 4: return (3 * e2 - 5) * x;
        case 5: return y;
        case 6: return (3 * e2 - 5) * y;
        case 7: return x * x - y * y;
        case 8: return x * y;
    }
    return 0;
}

double moment3(int r, int x, int y, int z) {
    const double e2 = x * x + y * y + z * z;
    switch (r) {
        case 0: return 1;
        case 1: return x;
        case 2: return y;
        case 3: return z;
        case 4: return x * y;
        case 5: return y * z;
        case 6: return x * z;
        case 7: return 3.0 * x * x - e2;
        case 8: return y * y - z * z;
        case 9: return e2 - 1;
        case 10: return (3 * e2 - 5) * x;
        case 11: return (3 * e2 - 5) * y;
        case 12: return (3 * e2 - 5) * z;
        case 13: return x * (y * y - z * z);
        case 14: return y * (z * z - x * x);
        case 15: return z * (x * x - y * y);
        case 16: return 3 * e2 * e2 - 6 * e2 + 1;
        case 17: return (2 * e2 - 3) * (3.0 * x * x - e2);
        case 18: return (2 * e2 - 3) * (y * y - z * z);
    }
    return 0;
}

void invert(int n, const double A[19][19], double out[19][19]) {
    long double a[19][38];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < 2 * n; ++j) a[i][j] = (j < n) ? A[i][j] : (j - n == i ? 1.0L : 0.0L);
    for (int c = 0; c < n; ++c) {
        int p = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs((double)a[r][c]) > std::fabs((double)a[p][c])) p = r;
        for (int j = 0; j < 2 * n; ++j) std::swap(a[c][j], a[p][j]);
        long double pv = a[c][c];
        for (int j = 0; j < 2 * n; ++j) a[c][j] /= pv;
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            long double f = a[r][c];
            if (f == 0) continue;
            for (int j = 0; j < 2 * n; ++j) a[r][j] -= f * a[c][j];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) out[i][j] = (double)a[i][n + j];
}

void make_lattice(int D, Lattice& L) {
    L.D = D;
    if (D == 2) {
        // WMRTvelSet, 2D branch (WMRT.cuh:192-196)
        static const int E[9][2] = {{0, 0}, {1, 0}, {0, 1}, {-1, 0}, {0, -1}, {1, 1}, {-1, 1}, {-1, -1}, {1, -1}};
        L.Q = 9;
        for (int q = 0; q < 9; ++q) {
            L.e[q][0] = E[q][0];
            L.e[q][1] = E[q][1];
            L.e[q][2] = 0;
            const int n2 = E[q][0] * E[q][0] + E[q][1] * E[q][1];
            L.w[q] = n2 == 0 ? 4.0 / 9.0 : (n2 == 1 ? 1.0 / 9.0 : 1.0 / 36.0);
        }
        L.nstress = 2;
        L.stress[0] = 7;
        L.stress[1] = 8;
        for (int r = 0; r < 9; ++r)
            for (int q = 0; q < 9; ++q) L.M[r][q] = moment2(r, L.e[q][0], L.e[q][1]);
    } else {
        // WMRTvelSet, 3D branch (WMRT.cuh:198-204)
        static const int E[19][3] = {{0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},
                                     {0, 0, 1},  {0, 0, -1}, {1, 1, 0},  {-1, -1, 0}, {1, -1, 0},
                                     {-1, 1, 0}, {1, 0, 1},  {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1},
                                     {0, 1, 1},  {0, -1, -1}, {0, 1, -1}, {0, -1, 1}};
        L.Q = 19;
        for (int q = 0; q < 19; ++q) {
            for (int d = 0; d < 3; ++d) L.e[q][d] = E[q][d];
            const int n2 = E[q][0] * E[q][0] + E[q][1] * E[q][1] + E[q][2] * E[q][2];
            L.w[q] = n2 == 0 ? 1.0 / 3.0 : (n2 == 1 ? 1.0 / 18.0 : 1.0 / 36.0);
        }
        L.nstress = 5;
        for (int s = 0; s < 5; ++s) L.stress[s] = 4 + s;
        for (int r = 0; r < 19; ++r)
            for (int q = 0; q < 19; ++q) L.M[r][q] = moment3(r, L.e[q][0], L.e[q][1], L.e[q][2]);
    }
    invert(L.Q, L.M, L.Minv);
    for (int q = 0; q < L.Q; ++q) {
        int den = 1;
        for (; den <= 1000; ++den) {
            bool ok = true;
            for (int r = 0; r < L.Q && ok; ++r) {
                const double x = L.Minv[q][r] * den;
                ok = std::fabs(x - std::nearbyint(x)) < 1e-9;
            }
            if (ok) break;
        }
        L.Minv_den[q] = den;
        for (int r = 0; r < L.Q; ++r) L.Minv_num[q][r] = std::nearbyint(L.Minv[q][r] * den);
    }
    for (int q = 0; q < L.Q; ++q)
        for (int p = 0; p < L.Q; ++p)
            if (L.e[p][0] == -L.e[q][0] && L.e[p][1] == -L.e[q][1] && L.e[p][2] == -L.e[q][2]) L.opp[q] = p;
    // side s = 2*axis + (0 minus | 1 plus); inward component +1 on a minus side
    L.nin = 0;
    for (int s = 0; s < 2 * D; ++s) {
        const int a = s / 2, sg = (s % 2 == 0) ? 1 : -1;
        int n = 0;
        for (int q = 0; q < L.Q; ++q) {  // normal first
            int on = 1;
            for (int d = 0; d < 3; ++d) on &= (d == a) ? (L.e[q][d] == sg) : (L.e[q][d] == 0);
            if (on) L.in[s][n++] = q;
        }
        for (int q = 0; q < L.Q; ++q) {
            if (L.e[q][a] != sg) continue;
            int normal = 1;
            for (int d = 0; d < 3; ++d)
                if (d != a && L.e[q][d] != 0) normal = 0;
            if (!normal) L.in[s][n++] = q;
        }
        L.nin = n;
    }
    L.weight = L.nin - 1;
}

// ---------------------------------------------------------------- state
struct Oracle {
    clip_config c;
    Lattice L;
    int D, Q;
    int ext[3], dmin[3], dmax[3], gmin[3], gmax[3];
    size_t n;
    std::vector<double> f, g, fpost, gpost, fprev, gprev, phi, rho, p, mu, dphi, normal, vel;
    bool isPeriodic = false, isWall = false, isSlip = false, isFC = false, isNeumann = false, isVelocity = false;
    long long step = 0;

    size_t idx(int i, int j, int k) const { return ((size_t)i * ext[1] + j) * ext[2] + k; }
    bool interior(int i, int j, int k) const {
        return i >= dmin[0] && i <= dmax[0] && j >= dmin[1] && j <= dmax[1] && k >= dmin[2] && k <= dmax[2];
    }
};

// Geometry::sdf (Geometry.cuh:102-161): first object whose id matches.
double sdf(const clip_config& c, int D, int id, double x, double y, double z) {
    for (int i = 0; i < c.num_geometries; ++i) {
        if (c.geom_id[i] != id) continue;
        const double* ce = c.geom_center[i];
        const double* le = c.geom_length[i];
        switch (c.geom_type[i]) {
            case CLIP_GEOM_CIRCLE: {
                double dx = x - ce[0], dy = y - ce[1];
                return std::sqrt(dx * dx + dy * dy) - c.geom_radius[i];
            }
            case CLIP_GEOM_SPHERE: {
                double dx = x - ce[0], dy = y - ce[1], dz = z - ce[2];
                return std::sqrt(dx * dx + dy * dy + dz * dz) - c.geom_radius[i];
            }
            case CLIP_GEOM_SQUARE: {
                double dx = std::fabs(x - ce[0]) - 0.5 * le[0];
                double dy = std::fabs(y - ce[1]) - 0.5 * le[1];
                double ax = dx > 0 ? dx : 0, ay = dy > 0 ? dy : 0;
                double m = dx > dy ? dx : dy;
                return std::sqrt(ax * ax + ay * ay) + (m < 0 ? m : 0);
            }
            case CLIP_GEOM_CUBE: {
                double dx = std::fabs(x - ce[0]) - 0.5 * le[0];
                double dy = std::fabs(y - ce[1]) - 0.5 * le[1];
                double dz = std::fabs(z - ce[2]) - 0.5 * le[2];
                double ax = dx > 0 ? dx : 0, ay = dy > 0 ? dy : 0, az = dz > 0 ? dz : 0;
                double m = dx > dy ? dx : dy;
                m = m > dz ? m : dz;
                return std::sqrt(ax * ax + ay * ay + az * az) + (m < 0 ? m : 0);
            }
            case CLIP_GEOM_PERTURBATION: {
                const double A = c.geom_amplitude[i] * le[0];
                double pert = (D == 2) ? A * std::cos(2.0 * M_PI * x / le[0])
                                       : A * (std::cos(2.0 * M_PI * x / le[0]) + std::cos(2.0 * M_PI * z / le[0]));
                return y - (pert + ce[1]);
            }
        }
    }
    return 1e10;
}

// Equilibrium_new (NsAllen.cu:77-94): velocity part of the equilibrium.
inline double gamma_u(const Lattice& L, int q, const double* u) {
    double eu = 0, u2 = 0;
    for (int d = 0; d < L.D; ++d) {
        eu += L.e[q][d] * u[d];
        u2 += u[d] * u[d];
    }
    return L.w[q] * (3.0 * eu + 4.5 * eu * eu - 1.5 * u2);
}

// The WMRT filter of NsAllen.cu:416-431 (m = M gneq, stress moments scaled
// by s, back-transform with M^-1), evaluated as the identity plus its
// stress-mode part: gneq + (s - 1) M^-1[:, stress] m[stress].  Same linear map
// (M^-1 M = I); only the stress rows of M and columns of M^-1 are touched.
#ifdef CLIP_ORACLE_LITERAL
// Built with -DCLIP_ORACLE_LITERAL (tests/test_oracle_order.py): the
// reference's literal operation order instead, m = M gneq over all rows (each
// row summed over q ascending, WMRT.cuh:54-126), the stress moments scaled by
// s (NsAllen.cu:428-429 / 448-452), then gneq = M^-1 m with M^-1 as integer
// numerators over the reference's per-row denominators (D2Q9: 36 for every
// row; D3Q19: 6, 36, 36, 72 x 4, 144 x 12; WMRT.cuh:77-92, 134-184), each row
// summed over the moments ascending and divided once.
void wmrt_filter(const Lattice& L, double* v, double s) {
    static const double den2[9] = {36, 36, 36, 36, 36, 36, 36, 36, 36};
    static const double den3[19] = {6, 36, 36, 72, 72, 72, 72, 144, 144, 144, 144, 144, 144, 144, 144, 144, 144, 144, 144};
    const double* den = (L.Q == 9) ? den2 : den3;
    auto accumulate = [](double acc, bool first, double c, double x) {
        if (first) return c == 1.0 ? x : (c == -1.0 ? -x : c * x);
        if (c == 1.0) return acc + x;
        if (c == -1.0) return acc - x;
        return acc + c * x;
    };
    double m[19];
    for (int r = 0; r < L.Q; ++r) {
        double acc = 0;
        bool first = true;
        for (int q = 0; q < L.Q; ++q) {
            if (L.M[r][q] == 0.0) continue;
            acc = accumulate(acc, first, L.M[r][q], v[q]);
            first = false;
        }
        m[r] = acc;
    }
    for (int k = 0; k < L.nstress; ++k) m[L.stress[k]] = m[L.stress[k]] * s;
    for (int q = 0; q < L.Q; ++q) {
        double acc = 0;
        bool first = true;
        for (int r = 0; r < L.Q; ++r) {
            const double c = std::nearbyint(L.Minv[q][r] * den[q]);
            if (c == 0.0) continue;
            acc = accumulate(acc, first, c, m[r]);
            first = false;
        }
        v[q] = acc / den[q];
    }
}
#else
void wmrt_filter(const Lattice& L, double* v, double s) {
    double m[5];
    for (int k = 0; k < L.nstress; ++k) {
        double acc = 0;
        for (int q = 0; q < L.Q; ++q) acc += L.M[L.stress[k]][q] * v[q];
        m[k] = acc;
    }
    const double sm1 = s - 1.0;
    for (int q = 0; q < L.Q; ++q) {
        double acc = 0;
        for (int k = 0; k < L.nstress; ++k) acc += L.Minv_num[q][L.stress[k]] * m[k];
        v[q] = v[q] + sm1 * (acc / L.Minv_den[q]);
    }
}
#endif

// tau(phi) with the clamp of NsAllen.cu:418-426.
inline double tau_of(const clip_config& c, double phi) {
    if (phi < 0.0) return c.tau_l;
    if (phi > 1.0) return c.tau_h;
    return c.tau_l + phi * (c.tau_h - c.tau_l);
}

// calculateVF (NsAllen.cu:102-132): viscous force from the filtered gneq,
// with the reference's term order for each component.
void viscous_force(const Oracle& o, const double* gneq, double tau, const double* dc, double* fv) {
    double sxx = 0, sxy = 0, syy = 0, szy = 0, szx = 0, szz = 0;
    for (int q = 0; q < o.Q; ++q) {
        const int* e = o.L.e[q];
        sxx += e[0] * e[0] * gneq[q];
        sxy += e[0] * e[1] * gneq[q];
        syy += e[1] * e[1] * gneq[q];
        if (o.D == 3) {
            szy += e[2] * e[1] * gneq[q];
            szx += e[2] * e[0] * gneq[q];
            szz += e[2] * e[2] * gneq[q];
        }
    }
    const double drho = o.c.rho_h - o.c.rho_l;
    const double dz = (o.D == 3) ? dc[2] : 0.0;
    fv[0] = -tau * (sxx * dc[0] + sxy * dc[1] + szx * dz) * drho;
    fv[1] = -tau * (sxy * dc[0] + syy * dc[1] + szy * dz) * drho;
    if (o.D == 3) fv[2] = -tau * (szz * dc[2] + szx * dc[0] + szy * dc[1]) * drho;
}

// Body force by case (NsAllen.cu:460-474).
void body_force(const Oracle& o, double rho, double* fg) {
    for (int d = 0; d < o.D; ++d) {
        const double gd = o.c.gravity[d];
        switch (o.c.case_type) {
            case CLIP_CASE_RTI: fg[d] = rho * gd; break;
            case CLIP_CASE_BUBBLE: fg[d] = (rho - o.c.rho_h) * gd; break;
            case CLIP_CASE_DROP: fg[d] = (rho - o.c.rho_l) * gd; break;
            default: fg[d] = 0.0; break;
        }
    }
}

// ------------------------------------------------------------ stencils
// Chemical_Potential (NsAllen.cu:207-236).
void chemical_potential(Oracle& o) {
#pragma omp parallel for schedule(static)
    for (int i = o.dmin[0]; i <= o.dmax[0]; ++i)
        for (int j = o.dmin[1]; j <= o.dmax[1]; ++j)
            for (int k = o.dmin[2]; k <= o.dmax[2]; ++k) {
                double lap;
                const double c0 = o.phi[o.idx(i, j, k)];
                if (o.D == 2) {
                    double corners = 0, edges = 0;
                    for (int a = -1; a <= 1; ++a)
                        for (int b = -1; b <= 1; ++b) {
                            const int m = std::abs(a) + std::abs(b);
                            const double v = o.phi[o.idx(i + a, j + b, 0)];
                            if (m == 2) corners += v;
                            if (m == 1) edges += v;
                        }
                    lap = (corners + 4.0 * edges - 20.0 * c0) / 6.0;
                } else {
                    double faces = 0, edges = 0, corners = 0;
                    for (int a = -1; a <= 1; ++a)
                        for (int b = -1; b <= 1; ++b)
                            for (int d = -1; d <= 1; ++d) {
                                const int m = std::abs(a) + std::abs(b) + std::abs(d);
                                const double v = o.phi[o.idx(i + a, j + b, k + d)];
                                if (m == 1) faces += v;
                                if (m == 2) edges += v;
                                if (m == 3) corners += v;
                            }
                    lap = (20.0 * faces + 6.0 * edges + corners - 200.0 * c0) / 48.0;
                }
                o.mu[o.idx(i, j, k)] = 4.0 * o.c.beta * c0 * (c0 - 1.0) * (c0 - 0.5) - o.c.kappa * lap;
            }
}

// Isotropic_Gradient (NsAllen.cu:280-322).
void isotropic_gradient(Oracle& o) {
#pragma omp parallel for schedule(static)
    for (int i = o.dmin[0]; i <= o.dmax[0]; ++i)
        for (int j = o.dmin[1]; j <= o.dmax[1]; ++j)
            for (int k = o.dmin[2]; k <= o.dmax[2]; ++k) {
                const size_t id = o.idx(i, j, k);
                for (int a = 0; a < o.D; ++a) {
                    // sum over neighbours with offset +1 / -1 along axis a,
                    // weighted by how many other axes they move along
                    double plus[3] = {0, 0, 0}, minus[3] = {0, 0, 0};
                    const int krange = (o.D == 3) ? 1 : 0;
                    for (int x = -1; x <= 1; ++x)
                        for (int y = -1; y <= 1; ++y)
                            for (int z = -krange; z <= krange; ++z) {
                                const int off[3] = {x, y, z};
                                if (off[a] == 0) continue;
                                int others = 0;
                                for (int b = 0; b < 3; ++b)
                                    if (b != a && off[b] != 0) others++;
                                const double v = o.phi[o.idx(i + x, j + y, k + z)];
                                if (off[a] > 0) plus[others] += v;
                                else minus[others] += v;
                            }
                    double gv;
                    if (o.D == 2) gv = (plus[0] - minus[0]) / 3.0 + (plus[1] - minus[1]) / 12.0;
                    else
                        gv = 0.5 * ((4.0 / 9.0) * (plus[0] - minus[0]) + (1.0 / 9.0) * (plus[1] - minus[1]) +
                                    (1.0 / 36.0) * (plus[2] - minus[2]));
                    o.dphi[id * o.D + a] = gv;
                }
            }
}

// normal_FD (NsAllen.cu:241-275).
void normals(Oracle& o) {
#pragma omp parallel for schedule(static)
    for (int i = o.dmin[0]; i <= o.dmax[0]; ++i)
        for (int j = o.dmin[1]; j <= o.dmax[1]; ++j)
            for (int k = o.dmin[2]; k <= o.dmax[2]; ++k) {
                const size_t id = o.idx(i, j, k);
                double s = 0;
                for (int a = 0; a < o.D; ++a) s += o.dphi[id * o.D + a] * o.dphi[id * o.D + a];
                const double t = MUTANT(8) ? s + 1e-32 : std::sqrt(s) + 1e-32;
                for (int a = 0; a < o.D; ++a)
                    o.normal[id * o.D + a] = o.dphi[id * o.D + (MUTANT(9) ? (a + 1) % o.D : a)] / t;
            }
}

// ------------------------------------------------------------ collision
// kernelCollisionMRTg (NsAllen.cu:545-584).
void collide_g(Oracle& o) {
    const double wc = 1.0 / (0.5 + 3.0 * o.c.mobility);
#pragma omp parallel for schedule(static)
    for (int i = o.dmin[0]; i <= o.dmax[0]; ++i)
        for (int j = o.dmin[1]; j <= o.dmax[1]; ++j)
            for (int k = o.dmin[2]; k <= o.dmax[2]; ++k) {
                const size_t id = o.idx(i, j, k);
                const double c0 = o.phi[id];
                const double* u = &o.vel[id * o.D];
                const double* nr = &o.normal[id * o.D];
                const double a = (3.0 * o.c.mobility * (1.0 - 4.0 * ((c0 - 0.5) * (c0 - 0.5)))) / o.c.interface_width;
                for (int q = 0; q < o.Q; ++q) {
                    double en = 0;
                    for (int d = 0; d < o.D; ++d) en += o.L.e[q][d] * nr[d];
                    const double heq = c0 * (gamma_u(o.L, q, u) + o.L.w[q]) + o.L.w[q] * (a * en);
                    o.g[id * o.Q + q] = o.gpost[id * o.Q + q] * (1.0 - wc) + heq * wc;
                }
            }
}

// kernelCollisionMRTf (NsAllen.cu:589-763).
void collide_f(Oracle& o) {
    const double drho3 = (o.c.rho_h - o.c.rho_l) / 3.0;
#pragma omp parallel for schedule(static)
    for (int i = o.dmin[0]; i <= o.dmax[0]; ++i)
        for (int j = o.dmin[1]; j <= o.dmax[1]; ++j)
            for (int k = o.dmin[2]; k <= o.dmax[2]; ++k) {
                const size_t id = o.idx(i, j, k);
                const double* fp = &o.fpost[id * o.Q];
                const double* u = &o.vel[id * o.D];
                const double* dc = &o.dphi[id * o.D];
                const double pp = o.p[id], rr = o.rho[id], c0 = o.phi[id], m0 = o.mu[id];
                double ga[19], gneq[19], hlp[19], fv[3] = {0, 0, 0}, fg[3] = {0, 0, 0};
                for (int q = 0; q < o.Q; ++q) {
                    ga[q] = gamma_u(o.L, q, u);
                    gneq[q] = fp[q] - (pp * o.L.w[q] + ga[q]);
                }
                const double tau = tau_of(o.c, c0);
                const double s = 1.0 / (tau + 0.5);
                wmrt_filter(o.L, gneq, s);
                viscous_force(o, gneq, tau, dc, fv);
                body_force(o, rr, fg);
                double F[3];
                for (int d = 0; d < o.D; ++d) F[d] = m0 * dc[d] + (-pp * drho3 * dc[d]) + fv[d] + fg[d];
                for (int q = 0; q < o.Q; ++q) {
                    double eF = 0;
                    for (int d = 0; d < o.D; ++d) eF += o.L.e[q][d] * F[d];
                    hlp[q] = 3.0 * o.L.w[q] * eF / rr;
                    gneq[q] = fp[q] - (pp * o.L.w[q] + ga[q] - 0.5 * hlp[q]);
                }
                wmrt_filter(o.L, gneq, s);
                for (int q = 0; q < o.Q; ++q) o.f[id * o.Q + q] = fp[q] - gneq[q] + hlp[q];
            }
}

// ------------------------------------------------------------ boundaries
bool axis_periodic(const Oracle& o, int a) {
    return o.c.bc_type[2 * a] == CLIP_BC_PERIODIC || o.c.bc_type[2 * a + 1] == CLIP_BC_PERIODIC;
}

// kernelPeriodicBoundary (Solver.cu:75-141): ghost <- opposite interior layer.
void periodic_pass(Oracle& o, std::vector<double>& a, int ndof) {
    for (int ax = 0; ax < o.D; ++ax) {
        if (!axis_periodic(o, ax)) continue;
        int lo[3] = {0, 0, 0}, hi[3] = {o.ext[0] - 1, o.ext[1] - 1, o.ext[2] - 1};
        lo[ax] = hi[ax] = 0;
#pragma omp parallel for schedule(static)
        for (int i = lo[0]; i <= hi[0]; ++i)
            for (int j = lo[1]; j <= hi[1]; ++j)
                for (int k = lo[2]; k <= hi[2]; ++k) {
                    int gl[3] = {i, j, k}, gh[3] = {i, j, k}, sl[3] = {i, j, k}, sh[3] = {i, j, k};
                    gl[ax] = o.gmin[ax];
                    gh[ax] = o.gmax[ax];
                    sl[ax] = o.dmax[ax];
                    sh[ax] = o.dmin[ax];
                    for (int q = 0; q < ndof; ++q) {
                        a[o.idx(gl[0], gl[1], gl[2]) * ndof + q] = a[o.idx(sl[0], sl[1], sl[2]) * ndof + q];
                        a[o.idx(gh[0], gh[1], gh[2]) * ndof + q] = a[o.idx(sh[0], sh[1], sh[2]) * ndof + q];
                    }
                }
    }
}

bool mirror_type(int t) {
    return t == CLIP_BC_WALL || t == CLIP_BC_FREE_CONVECT || t == CLIP_BC_NEUMANN || t == CLIP_BC_SLIP_WALL ||
           t == CLIP_BC_DO_NOTHING;
}

// Solver::mirrorBoundary + kernelMirrorBoundary (Solver.cu:927-978, 1048-1052).
void mirror_pass(Oracle& o) {
    if (!(o.isWall || o.isFC || o.isSlip || o.isNeumann)) return;
    for (int s = 0; s < 2 * o.D; ++s) {
        if (!mirror_type(o.c.bc_type[s])) continue;
        const int ax = s / 2;
        int lo[3] = {0, 0, 0}, hi[3] = {o.ext[0] - 1, o.ext[1] - 1, o.ext[2] - 1};
        lo[ax] = hi[ax] = 0;
#pragma omp parallel for schedule(static)
        for (int i = lo[0]; i <= hi[0]; ++i)
            for (int j = lo[1]; j <= hi[1]; ++j)
                for (int k = lo[2]; k <= hi[2]; ++k) {
                    int gc[3] = {i, j, k}, bc[3] = {i, j, k};
                    gc[ax] = (s % 2 == 0) ? o.gmin[ax] : o.gmax[ax];
                    bc[ax] = (s % 2 == 0) ? o.dmin[ax] : o.dmax[ax];
                    o.phi[o.idx(gc[0], gc[1], gc[2])] = o.phi[o.idx(bc[0], bc[1], bc[2])];
                }
    }
}

// Inlet velocity at a ghost position of side s (its `value`, optionally
// modulated by azimuthal modes around geometry 0's centre).
void inlet_velocity(const Oracle& o, int s, const int* pos, double* U) {
    for (int d = 0; d < 3; ++d) U[d] = o.c.bc_value[s][d];
    if (o.c.inlet_modes <= 0 || o.c.num_geometries <= 0) return;
    const int ax = s / 2;
    int t1 = (ax + 1) % 3, t2 = (ax + 2) % 3;
    if (o.D == 2) {
        t1 = 1 - ax;
        t2 = -1;
    }
    const double d1 = pos[t1] - o.c.geom_center[0][t1];
    const double d2 = (t2 >= 0) ? pos[t2] - o.c.geom_center[0][t2] : 0.0;
    const double th = std::atan2(d2, d1);
    double mod = 1.0;
    for (int m = 0; m < o.c.inlet_modes; ++m) mod += o.c.inlet_mode_amp[m] * std::cos((m + 1) * th + o.c.inlet_mode_phase[m]);
    for (int d = 0; d < 3; ++d) U[d] *= mod;
}

// JetBoundary (Solver.cu:550-921), one side after another.
void velocity_pass(Oracle& o) {
    if (!o.isVelocity) return;
    for (int s = 0; s < 2 * o.D; ++s) {
        if (o.c.bc_type[s] != CLIP_BC_VELOCITY) continue;
        const int ax = s / 2;
        const int* in = o.L.in[s];
        int lo[3] = {0, 0, 0}, hi[3] = {o.ext[0] - 1, o.ext[1] - 1, o.ext[2] - 1};
        lo[ax] = hi[ax] = 0;
#pragma omp parallel for schedule(static)
        for (int i = lo[0]; i <= hi[0]; ++i)
            for (int j = lo[1]; j <= hi[1]; ++j)
                for (int k = lo[2]; k <= hi[2]; ++k) {
                    int gc[3] = {i, j, k};
                    gc[ax] = (s % 2 == 0) ? o.gmin[ax] : o.gmax[ax];
                    const size_t gid = o.idx(gc[0], gc[1], gc[2]);
                    const double z = (o.D == 3) ? (double)gc[2] : 0.0;
                    if (sdf(o.c, o.D, 0, gc[0], gc[1], z) <= 0) {
                        double U[3];
                        inlet_velocity(o, s, gc, U);
                        double feq[19], geq[19], Mt[3] = {0, 0, 0}, N = 0;
                        const double cg = o.phi[gid];
                        for (int q = 1; q < o.Q; ++q) {
                            const double fa = gamma_u(o.L, q, U);
                            feq[q] = fa;
                            geq[q] = cg * (fa + o.L.w[q]);
                            if (o.L.e[q][ax] == 0) {
                                for (int t = 0; t < o.D; ++t)
                                    if (t != ax) Mt[t] += o.L.e[q][t] * (o.f[gid * o.Q + q] - feq[q]);
                                N += o.g[gid * o.Q + q] - geq[q];
                            }
                        }
                        for (int m = 0; m < o.L.nin; ++m) {
                            const int a = in[m], b = o.L.opp[a];
                            double corr = 0;
                            for (int t = 0; t < o.D; ++t)
                                if (t != ax) corr += o.L.e[a][t] * Mt[t];
                            const double wt = MUTANT(2) ? o.L.weight + 1.0 : o.L.weight;
                            const double sg = MUTANT(3) ? -1.0 : 1.0;
                            o.f[gid * o.Q + a] = MUTANT(4) ? o.f[gid * o.Q + b] + feq[b] - feq[a] - (1.0 / wt) * corr
                                                           : o.f[gid * o.Q + b] + feq[a] - feq[b] - sg * (1.0 / wt) * corr;
                            double gv = geq[a] - (o.g[gid * o.Q + b] - geq[b]);
                            if (m != 0 || MUTANT(1)) gv -= N / o.L.weight;
                            o.g[gid * o.Q + a] = gv;
                        }
                    } else {
                        int bc[3] = {gc[0], gc[1], gc[2]};
                        bc[ax] = (s % 2 == 0) ? o.dmin[ax] : o.dmax[ax];
                        const size_t bid = o.idx(bc[0], bc[1], bc[2]);
                        for (int m = 0; m < o.L.nin; ++m) {
                            const int src = MUTANT(5) ? in[m] : o.L.opp[in[m]];
                            o.g[gid * o.Q + in[m]] = o.g[bid * o.Q + src];
                            o.f[gid * o.Q + in[m]] = o.f[bid * o.Q + src];
                        }
                        o.phi[gid] = o.phi[bid];
                    }
                }
    }
}

// kernelStreaming (NsAllen.cu:327-364) after the f_prev copy (847-854).
void streaming(Oracle& o) {
    if (o.isFC) {
        o.fprev = o.fpost;
        o.gprev = o.gpost;
    }
#pragma omp parallel for schedule(static)
    for (int i = o.dmin[0]; i <= o.dmax[0]; ++i)
        for (int j = o.dmin[1]; j <= o.dmax[1]; ++j)
            for (int k = o.dmin[2]; k <= o.dmax[2]; ++k) {
                const size_t id = o.idx(i, j, k);
                for (int q = 0; q < o.Q; ++q) {
                    const size_t src = o.idx(i - o.L.e[q][0], j - o.L.e[q][1], k - o.L.e[q][2]);
                    o.fpost[id * o.Q + q] = o.f[src * o.Q + q];
                    o.gpost[id * o.Q + q] = o.g[src * o.Q + q];
                }
            }
}

bool on_side(const Oracle& o, int s, int i, int j, int k) {
    const int c[3] = {i, j, k};
    const int ax = s / 2;
    return c[ax] == ((s % 2 == 0) ? o.dmin[ax] : o.dmax[ax]);
}

// The reference's boundary maps as it lists them, sides X-, X+, Y-, Y+, Z-,
// Z+: WMRT::wallBCMap (WMRT.cuh:214-237) and WMRT::slipWallBCMap
// (WMRT.cuh:243-264).  kernelHalfBounceBack assigns a_post[b, map[s][m]] =
// a[b, map[s^1][m]] for list position m outer and side s inner.
const int kWallMap2[4][3] = {{1, 5, 8}, {3, 7, 6}, {2, 5, 6}, {4, 7, 8}};
const int kSlipMap2[4][3] = {{1, 5, 8}, {3, 6, 7}, {2, 5, 6}, {4, 8, 7}};
const int kWallMap3[6][5] = {{1, 9, 7, 13, 11}, {2, 10, 8, 14, 12}, {3, 10, 7, 17, 15},
                             {4, 9, 8, 18, 16}, {5, 14, 11, 18, 15}, {6, 13, 12, 17, 16}};
const int kSlipMap3[6][5] = {{1, 9, 7, 13, 11}, {2, 8, 10, 12, 14}, {3, 10, 7, 17, 15},
                             {4, 8, 9, 16, 18}, {5, 14, 11, 18, 15}, {6, 12, 13, 16, 17}};

int side_map(const Oracle& o, bool slip, int s, int m) {
    if (o.D == 2) return slip ? kSlipMap2[s][m] : kWallMap2[s][m];
    return slip ? kSlipMap3[s][m] : kWallMap3[s][m];
}

// kernelHalfBounceBack with the slip map applied to slip sides and the wall
// map to wall sides, in the reference kernel's loop order (the opt-in
// specular slip wall, clip_config::slip_specular).  Launched when any side
// is a wall or a slip wall.
void wall_pass_specular(Oracle& o) {
    if (!(o.isWall || o.isSlip)) return;
    const int nin = (o.D == 2) ? 3 : 5;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o.ext[0]; ++i)
        for (int j = 0; j < o.ext[1]; ++j)
            for (int k = 0; k < o.ext[2]; ++k) {
                const size_t id = o.idx(i, j, k);
                if (!o.interior(i, j, k)) continue;
                for (int m = 0; m < nin; ++m)
                    for (int s = 0; s < 2 * o.D; ++s) {
                        const int t = o.c.bc_type[s];
                        if (!(t == CLIP_BC_WALL || t == CLIP_BC_SLIP_WALL) || !on_side(o, s, i, j, k)) continue;
                        const bool slip = MUTANT(10) ? false : t == CLIP_BC_SLIP_WALL;
                        const int q = side_map(o, slip, s, m);
                        const int b = MUTANT(11) ? side_map(o, slip, s ^ 1, (m + 1) % nin) : side_map(o, slip, s ^ 1, m);
                        o.fpost[id * o.Q + q] = o.f[id * o.Q + b];
                        o.gpost[id * o.Q + q] = o.g[id * o.Q + b];
                    }
            }
}

// Solver::slipWallBoundary (Solver.cu:1012-1016) as the reference defines it:
// kernelHalfBounceBack with slipWallBCMap over every Wall and SlipWall side,
// launched when isSlipWall.  NSAllen::solve never calls it; clip_amd exposes
// it as a launcher on caller arrays.
void slip_wall_launcher(Oracle& o) {
    if (!o.isSlip) return;
    const int nin = (o.D == 2) ? 3 : 5;
    for (int i = 0; i < o.ext[0]; ++i)
        for (int j = 0; j < o.ext[1]; ++j)
            for (int k = 0; k < o.ext[2]; ++k) {
                if (!o.interior(i, j, k)) continue;
                const size_t id = o.idx(i, j, k);
                for (int m = 0; m < nin; ++m)
                    for (int s = 0; s < 2 * o.D; ++s) {
                        const int t = o.c.bc_type[s];
                        if (!(t == CLIP_BC_WALL || t == CLIP_BC_SLIP_WALL) || !on_side(o, s, i, j, k)) continue;
                        const int q = side_map(o, true, s, m), b = side_map(o, true, s ^ 1, m);
                        o.fpost[id * o.Q + q] = o.f[id * o.Q + b];
                        o.gpost[id * o.Q + q] = o.g[id * o.Q + b];
                    }
            }
}

// kernelHalfBounceBack with wallBCMap (Solver.cu:248-333, launched when isWall).
void wall_pass(Oracle& o) {
    if (o.c.slip_specular) return wall_pass_specular(o);
    if (!o.isWall) return;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o.ext[0]; ++i)
        for (int j = 0; j < o.ext[1]; ++j)
            for (int k = 0; k < o.ext[2]; ++k) {
                const size_t id = o.idx(i, j, k);
                for (int s = 0; s < 2 * o.D; ++s) {
                    const int t = o.c.bc_type[s];
                    if (!(t == CLIP_BC_WALL || t == CLIP_BC_SLIP_WALL) || !on_side(o, s, i, j, k)) continue;
                    for (int m = 0; m < o.L.nin; ++m) {
                        const int q = o.L.in[s][m], b = o.L.opp[q];
                        o.fpost[id * o.Q + q] = o.f[id * o.Q + b];
                        o.gpost[id * o.Q + q] = o.g[id * o.Q + b];
                    }
                }
            }
}

// kernelFreeConvect (Solver.cu:342-453).
void free_convect_pass(Oracle& o) {
    if (!o.isFC) return;
    const std::vector<double> f0 = o.fpost, g0 = o.gpost;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o.ext[0]; ++i)
        for (int j = 0; j < o.ext[1]; ++j)
            for (int k = 0; k < o.ext[2]; ++k) {
                const size_t id = o.idx(i, j, k);
                for (int s = 0; s < 2 * o.D; ++s) {
                    if (o.c.bc_type[s] != CLIP_BC_FREE_CONVECT || !on_side(o, s, i, j, k)) continue;
                    const int ax = s / 2, sg = (s % 2 == 0) ? 1 : -1;
                    int nc[3] = {i, j, k};
                    nc[ax] += sg;
                    const size_t nid = o.idx(nc[0], nc[1], nc[2]);
                    const double cv = std::fabs(o.vel[nid * o.D + ax]);
                    for (int m = 0; m < o.L.nin; ++m) {
                        const int q = o.L.in[s][m];
                        o.fpost[id * o.Q + q] = (o.fprev[id * o.Q + q] + cv * f0[nid * o.Q + q]) / (1.0 + cv);
                        o.gpost[id * o.Q + q] = (o.gprev[id * o.Q + q] + cv * g0[nid * o.Q + q]) / (1.0 + cv);
                    }
                }
            }
}

// kernelNeumann (Solver.cu:462-542).
void neumann_pass(Oracle& o) {
    if (!o.isNeumann) return;
    const std::vector<double> f0 = o.fpost, g0 = o.gpost;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o.ext[0]; ++i)
        for (int j = 0; j < o.ext[1]; ++j)
            for (int k = 0; k < o.ext[2]; ++k) {
                const size_t id = o.idx(i, j, k);
                for (int s = 0; s < 2 * o.D; ++s) {
                    if (o.c.bc_type[s] != CLIP_BC_NEUMANN || !on_side(o, s, i, j, k)) continue;
                    const int ax = s / 2, sg = (s % 2 == 0) ? 1 : -1;
                    int nc[3] = {i, j, k};
                    nc[ax] += sg;
                    const size_t nid = o.idx(nc[0], nc[1], nc[2]);
                    for (int m = 0; m < o.L.nin; ++m) {
                        const int q = o.L.in[s][m];
                        o.fpost[id * o.Q + q] = f0[nid * o.Q + q];
                        o.gpost[id * o.Q + q] = g0[nid * o.Q + q];
                    }
                }
            }
}

// ------------------------------------------------------------ macroscopic
// kernelMacroscopicg (NsAllen.cu:516-540).
void macro_g(Oracle& o) {
#pragma omp parallel for schedule(static)
    for (int i = o.dmin[0]; i <= o.dmax[0]; ++i)
        for (int j = o.dmin[1]; j <= o.dmax[1]; ++j)
            for (int k = o.dmin[2]; k <= o.dmax[2]; ++k) {
                const size_t id = o.idx(i, j, k);
                double c0 = 0;
                for (int q = 0; q < o.Q; ++q) c0 += o.gpost[id * o.Q + q];
                o.phi[id] = c0;
                o.rho[id] = o.c.rho_l + c0 * (o.c.rho_h - o.c.rho_l);
            }
}

// kernelMacroscopicf (NsAllen.cu:369-511).
void macro_f(Oracle& o) {
    const double drho3 = (o.c.rho_h - o.c.rho_l) / 3.0;
#pragma omp parallel for schedule(static)
    for (int i = o.dmin[0]; i <= o.dmax[0]; ++i)
        for (int j = o.dmin[1]; j <= o.dmax[1]; ++j)
            for (int k = o.dmin[2]; k <= o.dmax[2]; ++k) {
                const size_t id = o.idx(i, j, k);
                const double* fp = &o.fpost[id * o.Q];
                double* u = &o.vel[id * o.D];
                const double* dc = &o.dphi[id * o.D];
                double pp = 0;
                for (int q = 0; q < o.Q; ++q) pp += fp[q];
                o.p[id] = pp;
                double gneq[19], fv[3] = {0, 0, 0}, fg[3] = {0, 0, 0};
                for (int q = 0; q < o.Q; ++q) gneq[q] = fp[q] - (pp * o.L.w[q] + gamma_u(o.L, q, u));
                const double c0 = o.phi[id], rr = o.rho[id];
                const double tau = tau_of(o.c, c0);
                wmrt_filter(o.L, gneq, 1.0 / (tau + 0.5));
                viscous_force(o, gneq, tau, dc, fv);
                body_force(o, rr, fg);
                double F[3];
                for (int d = 0; d < o.D; ++d) F[d] = o.mu[id] * dc[d] + (-pp * drho3 * dc[d]) + fv[d] + fg[d];
                double mom[3] = {0, 0, 0};
                for (int q = 0; q < o.Q; ++q)
                    for (int d = 0; d < o.D; ++d) mom[d] += o.L.e[q][d] * fp[q];
                for (int d = 0; d < o.D; ++d) u[d] = mom[d] + 0.5 * F[d] / rr;
            }
}

void macroscopic(Oracle& o) {
    macro_g(o);
    chemical_potential(o);
    isotropic_gradient(o);
    macro_f(o);
}

// NSAllen::solve (NsAllen.cu:881-892).
void solve(Oracle& o) {
    normals(o);
    collide_g(o);
    collide_f(o);
    if (o.isPeriodic) {
        periodic_pass(o, o.f, o.Q);
        periodic_pass(o, o.g, o.Q);
        periodic_pass(o, o.phi, 1);
    }
    velocity_pass(o);
    streaming(o);
    wall_pass(o);
    mirror_pass(o);
    free_convect_pass(o);
    mirror_pass(o);
    neumann_pass(o);
    mirror_pass(o);
    macroscopic(o);
    o.step++;
}

// NSAllen::initialCondition (NsAllen.cu:768-824), host-side phi.
void initial_phi(Oracle& o) {
    for (int i = 0; i < o.ext[0]; ++i)
        for (int j = 0; j < o.ext[1]; ++j)
            for (int k = 0; k < o.ext[2]; ++k) {
                const double x = i, y = j, z = (o.D == 3) ? (double)k : 0.0;
                const double W = o.c.interface_width;
                double v = 0.0;
                switch (o.c.case_type) {
                    case CLIP_CASE_BUBBLE: v = 0.5 + 0.5 * std::tanh(2.0 * sdf(o.c, o.D, 0, x, y, z) / W); break;
                    case CLIP_CASE_DROP:
                    case CLIP_CASE_JET: v = 0.5 - 0.5 * std::tanh(2.0 * sdf(o.c, o.D, 0, x, y, z) / W); break;
                    case CLIP_CASE_RTI: v = 0.5 + 0.5 * std::tanh(2.0 * sdf(o.c, o.D, 0, x, y, z) / W); break;
                    default: v = 0.0;
                }
                o.phi[o.idx(i, j, k)] = v;
            }
}

// NSAllen::deviceInitializer (NsAllen.cu:897-908) with KernelInitializeDistributions (137-202).
void device_initializer(Oracle& o) {
    chemical_potential(o);
    isotropic_gradient(o);
    normals(o);
    const double L2 = o.c.reference_length / 2.0;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o.ext[0]; ++i)
        for (int j = 0; j < o.ext[1]; ++j)
            for (int k = 0; k < o.ext[2]; ++k) {
                const size_t id = o.idx(i, j, k);
                const double c0 = o.phi[id];
                const double rr = o.c.rho_l + c0 * (o.c.rho_h - o.c.rho_l);
                o.rho[id] = rr;
                double pp = 0.0;
                if (o.c.case_type == CLIP_CASE_BUBBLE) pp = 0.0 - c0 * o.c.sigma / L2 / (rr / 3.0);
                else if (o.c.case_type == CLIP_CASE_DROP) pp = 0.0 + c0 * o.c.sigma / L2 / (rr / 3.0);
                o.p[id] = pp;
                double* u = &o.vel[id * o.D];
                for (int d = 0; d < o.D; ++d) u[d] = 0.0;
                const double* nr = &o.normal[id * o.D];
                for (int q = 0; q < o.Q; ++q) {
                    const double ga = gamma_u(o.L, q, u);
                    double en = 0;
                    for (int d = 0; d < o.D; ++d) en += o.L.e[q][d] * nr[d];
                    const double hlp = o.L.w[q] * ((1.0 - 4.0 * ((c0 - 0.5) * (c0 - 0.5))) / o.c.interface_width * en);
                    const double gv = MUTANT(6)   ? c0 * (ga + o.L.w[q]) + 0.5 * hlp
                                      : MUTANT(7) ? c0 * (ga + o.L.w[q]) - 0.5 * 3.0 * o.c.mobility * hlp
                                                  : c0 * (ga + o.L.w[q]) - 0.5 * hlp;
                    const double fv = pp * o.L.w[q] + ga;
                    o.g[id * o.Q + q] = o.gpost[id * o.Q + q] = gv;
                    o.f[id * o.Q + q] = o.fpost[id * o.Q + q] = fv;
                }
            }
    if (o.isFC) {
        o.fprev = o.fpost;
        o.gprev = o.gpost;
    }
}

std::vector<double>* field_ptr(Oracle& o, int field, int& ndof) {
    switch (field) {
        case CLIP_FIELD_PHI: ndof = 1; return &o.phi;
        case CLIP_FIELD_RHO: ndof = 1; return &o.rho;
        case CLIP_FIELD_P: ndof = 1; return &o.p;
        case CLIP_FIELD_U: ndof = o.D; return &o.vel;
        case CLIP_FIELD_MU: ndof = 1; return &o.mu;
        case CLIP_FIELD_DPHI: ndof = o.D; return &o.dphi;
        case CLIP_FIELD_NORMAL: ndof = o.D; return &o.normal;
        case CLIP_FIELD_F: ndof = o.Q; return &o.f;
        case CLIP_FIELD_G: ndof = o.Q; return &o.g;
        case CLIP_FIELD_F_POST: ndof = o.Q; return &o.fpost;
        case CLIP_FIELD_G_POST: ndof = o.Q; return &o.gpost;
    }
    return nullptr;
}

}  // namespace

extern "C" {

void* orc_create(const clip_config* cfg) {
    if (!cfg || (cfg->dim != 2 && cfg->dim != 3)) return nullptr;
    Oracle* o = new Oracle();
    o->c = *cfg;
    o->D = cfg->dim;
    make_lattice(o->D, o->L);
    o->Q = o->L.Q;
    if (o->D == 2) {
        o->ext[0] = (int)cfg->N[0] + 2;
        o->ext[1] = (int)cfg->N[1] + 2;
        o->ext[2] = 1;
        for (int a = 0; a < 2; ++a) {
            o->dmin[a] = 1;
            o->dmax[a] = o->ext[a] - 2;
            o->gmin[a] = 0;
            o->gmax[a] = o->ext[a] - 1;
        }
        o->dmin[2] = o->dmax[2] = o->gmin[2] = o->gmax[2] = 0;
    } else {
        for (int a = 0; a < 3; ++a) {
            o->ext[a] = (int)cfg->N[a];
            o->dmin[a] = 1;
            o->dmax[a] = o->ext[a] - 2;
            o->gmin[a] = 0;
            o->gmax[a] = o->ext[a] - 1;
        }
    }
    o->n = (size_t)o->ext[0] * o->ext[1] * o->ext[2];
    const size_t n = o->n, Q = o->Q, D = o->D;
    o->f.assign(n * Q, 0);
    o->g.assign(n * Q, 0);
    o->fpost.assign(n * Q, 0);
    o->gpost.assign(n * Q, 0);
    o->phi.assign(n, 0);
    o->rho.assign(n, 0);
    o->p.assign(n, 0);
    o->mu.assign(n, 0);
    o->dphi.assign(n * D, 0);
    o->normal.assign(n * D, 0);
    o->vel.assign(n * D, 0);
    for (int s = 0; s < 2 * o->D; ++s) {
        switch (cfg->bc_type[s]) {
            case CLIP_BC_WALL: o->isWall = true; break;
            case CLIP_BC_SLIP_WALL: o->isSlip = true; break;
            case CLIP_BC_FREE_CONVECT: o->isFC = true; break;
            case CLIP_BC_NEUMANN: o->isNeumann = true; break;
            case CLIP_BC_PERIODIC: o->isPeriodic = true; break;
            case CLIP_BC_VELOCITY: o->isVelocity = true; break;
        }
    }
    if (o->isFC) {
        o->fprev.assign(n * Q, 0);
        o->gprev.assign(n * Q, 0);
    }
    return o;
}

int orc_init(void* h) {
    Oracle* o = (Oracle*)h;
    initial_phi(*o);
    device_initializer(*o);
    o->step = 0;
    return 0;
}

// Same, with phi given on every node (CLIP AoS order, ghosts included)
// instead of the case's initial condition.
int orc_init_phi(void* h, const double* phi) {
    Oracle* o = (Oracle*)h;
    for (size_t i = 0; i < o->phi.size(); ++i) o->phi[i] = phi[i];
    device_initializer(*o);
    o->step = 0;
    return 0;
}

int orc_step(void* h, int n) {
    Oracle* o = (Oracle*)h;
    for (int s = 0; s < n; ++s) solve(*o);
    return 0;
}

// Only the collision phase (normals, g and f collision) — used to time and
// test one phase in isolation.
int orc_collide(void* h) {
    Oracle* o = (Oracle*)h;
    normals(*o);
    collide_g(*o);
    collide_f(*o);
    return 0;
}

int orc_macroscopic(void* h) {
    macroscopic(*(Oracle*)h);
    return 0;
}

// One named phase of the step, for the invariant tests:
// "normals", "collide_g", "collide_f", "periodic", "velocity", "stream",
// "wall", "mirror", "free_convect", "neumann", "macro_g", "chem_pot",
// "gradient", "macro_f".
int orc_phase(void* h, const char* name) {
    Oracle& o = *(Oracle*)h;
    const std::string n(name);
    if (n == "normals") normals(o);
    else if (n == "collide_g") collide_g(o);
    else if (n == "collide_f") collide_f(o);
    else if (n == "periodic") {
        periodic_pass(o, o.f, o.Q);
        periodic_pass(o, o.g, o.Q);
        periodic_pass(o, o.phi, 1);
    } else if (n == "velocity") velocity_pass(o);
    else if (n == "stream") streaming(o);
    else if (n == "wall") wall_pass(o);
    else if (n == "slip_wall") slip_wall_launcher(o);
    else if (n == "mirror") mirror_pass(o);
    else if (n == "free_convect") free_convect_pass(o);
    else if (n == "neumann") neumann_pass(o);
    else if (n == "macro_g") macro_g(o);
    else if (n == "chem_pot") chemical_potential(o);
    else if (n == "gradient") isotropic_gradient(o);
    else if (n == "macro_f") macro_f(o);
    else return -1;
    return 0;
}

// f_prev / g_prev (outflow) of the oracle, same shape as F.
int orc_get_prev(void* h, int which, double* out) {
    Oracle& o = *(Oracle*)h;
    const std::vector<double>& v = which ? o.gprev : o.fprev;
    if (v.empty()) return -1;
    std::memcpy(out, v.data(), v.size() * sizeof(double));
    return 0;
}
int orc_set_prev(void* h, int which, const double* in) {
    Oracle& o = *(Oracle*)h;
    std::vector<double>& v = which ? o.gprev : o.fprev;
    if (v.empty()) return -1;
    std::memcpy(v.data(), in, v.size() * sizeof(double));
    return 0;
}

// Velocity part of the equilibrium for one direction (Equilibrium_new).
double orc_gamma(int dim, int q, const double* u) {
    Lattice L;
    make_lattice(dim, L);
    return gamma_u(L, q, u);
}

long long orc_field_size(void* h, int field) {
    Oracle* o = (Oracle*)h;
    int nd = 0;
    std::vector<double>* v = field_ptr(*o, field, nd);
    return v ? (long long)(o->n * nd) : -1;
}

int orc_get(void* h, int field, double* out) {
    Oracle* o = (Oracle*)h;
    int nd = 0;
    std::vector<double>* v = field_ptr(*o, field, nd);
    if (!v) return -1;
    std::memcpy(out, v->data(), v->size() * sizeof(double));
    return 0;
}

int orc_set(void* h, int field, const double* in) {
    Oracle* o = (Oracle*)h;
    int nd = 0;
    std::vector<double>* v = field_ptr(*o, field, nd);
    if (!v) return -1;
    std::memcpy(v->data(), in, v->size() * sizeof(double));
    return 0;
}

// Lattice tables as the oracle built them (for the invariant tests).
int orc_lattice(int dim, double* M, double* Minv, int* e, double* w, int* in_lists, int* nin) {
    Lattice L;
    make_lattice(dim, L);
    for (int r = 0; r < L.Q; ++r)
        for (int q = 0; q < L.Q; ++q) {
            M[r * L.Q + q] = L.M[r][q];
            Minv[r * L.Q + q] = L.Minv[r][q];
        }
    for (int q = 0; q < L.Q; ++q) {
        for (int d = 0; d < 3; ++d) e[q * 3 + d] = L.e[q][d];
        w[q] = L.w[q];
    }
    for (int s = 0; s < 2 * dim; ++s)
        for (int m = 0; m < L.nin; ++m) in_lists[s * L.nin + m] = L.in[s][m];
    *nin = L.nin;
    return L.Q;
}

// The reference's wall (slip = 0) or slip (slip = 1) map, 2*dim lists of
// nin directions (the oracle's literal tables).
int orc_side_map(int dim, int slip, int* out) {
    const int nin = (dim == 2) ? 3 : 5;
    for (int s = 0; s < 2 * dim; ++s)
        for (int m = 0; m < nin; ++m)
            out[s * nin + m] = (dim == 2) ? (slip ? kSlipMap2[s][m] : kWallMap2[s][m])
                                          : (slip ? kSlipMap3[s][m] : kWallMap3[s][m]);
    return nin;
}

// Single-node kernels exposed for the invariant tests.
int orc_filter(int dim, double* v, double s) {
    Lattice L;
    make_lattice(dim, L);
    wmrt_filter(L, v, s);
    return 0;
}

double orc_sdf(const clip_config* c, double x, double y, double z) { return sdf(*c, c->dim, 0, x, y, z); }

// calculateVF of one node with the handle's lattice and densities.
int orc_viscous_force(void* h, const double* gneq, double tau, const double* dc, double* fv) {
    viscous_force(*(Oracle*)h, gneq, tau, dc, fv);
    return 0;
}

void orc_destroy(void* h) { delete (Oracle*)h; }

}  // extern "C"
