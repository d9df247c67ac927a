// node_check.hip — host build of the per-node arithmetic (node.hpp) for the
// CPU tests: forward node update and its reverse, FP64, on D2Q9 and D3Q19.
// phys = [rho_h, rho_l, tau_h, tau_l, beta, kappa, mobility, W, g0, g1, g2, case_type].
#include "node.hpp"

using namespace clipamd;

namespace {

template <class T = double>
Phys<T> unpack(const double* a) {
    Phys<T> ph;
    ph.rho_h = a[0];
    ph.rho_l = a[1];
    ph.tau_h = a[2];
    ph.tau_l = a[3];
    ph.beta = a[4];
    ph.kappa = a[5];
    ph.mobility = a[6];
    ph.W = a[7];
    ph.grav[0] = a[8];
    ph.grav[1] = a[9];
    ph.grav[2] = a[10];
    ph.case_type = (int)a[11];
    return ph;
}

// FP32 reverse of one node (inputs given in FP64, rounded to FP32).
template <class L>
void rev32(const double* phys, const double* fp, const double* gp, double c0, const double* dc, double lap,
           const double* uo, const double* lfs, const double* lgs, const double* lu, double nfloor, double* lfp,
           double* lgp, double* lin) {
    const Phys<float> ph = unpack<float>(phys);
    float f[L::Q], g[L::Q], a[L::Q], b[L::Q];
    for (int q = 0; q < L::Q; ++q) {
        f[q] = (float)fp[q];
        g[q] = (float)gp[q];
        a[q] = (float)lfs[q];
        b[q] = (float)lgs[q];
    }
    float dcc[3] = {(float)dc[0], (float)dc[1], L::D == 3 ? (float)dc[2] : 0.f};
    float uoo[3] = {(float)uo[0], (float)uo[1], L::D == 3 ? (float)uo[2] : 0.f};
    float luu[3] = {(float)lu[0], (float)lu[1], L::D == 3 ? (float)lu[2] : 0.f};
    NodeAdj<L, float> o{};
    node_reverse<L, float>(ph, f, g, (float)c0, dcc, (float)lap, uoo, a, b, luu, o, (float)nfloor);
    for (int q = 0; q < L::Q; ++q) {
        lfp[q] = o.lfp[q];
        lgp[q] = o.lgp[q];
    }
    lin[0] = o.lc0;
    for (int d = 0; d < 3; ++d) lin[1 + d] = o.ldc[d];
    lin[4] = o.llap;
    for (int d = 0; d < 3; ++d) lin[5 + d] = o.luo[d];
    lin[8] = o.lbeta;
    lin[9] = o.lkappa;
    lin[10] = o.lmob;
}

template <class L>
void fwd(const double* phys, const double* fp, const double* gp, double c0, const double* dc, double lap,
         const double* uo, double* fs, double* gs, double* u) {
    const Phys<double> ph = unpack(phys);
    double p = 0;
    for (int q = 0; q < L::Q; ++q) p += fp[q];
    double mom[3], Pf[3][3];
    pop_moments<L>(fp, mom, Pf);
    const NodeScalars<double> z = node_scalars<L>(ph, c0, lap);
    double dcc[3] = {dc[0], dc[1], L::D == 3 ? dc[2] : 0.0};
    double uoo[3] = {uo[0], uo[1], L::D == 3 ? uo[2] : 0.0};
    double uu[3];
    velocity_of<L>(z, p, mom, Pf, dcc, uoo, uu);
    for (int d = 0; d < L::D; ++d) u[d] = uu[d];
    collide_node<L, double>(
        z, p, Pf, c0, dcc, uu, gp, [&](int q, double v) { gs[q] = v; }, [&](int q, double v) { fs[q] = v; });
}

template <class L>
void rev(const double* phys, const double* fp, const double* gp, double c0, const double* dc, double lap,
         const double* uo, const double* lfs, const double* lgs, const double* lu, double* lfp, double* lgp,
         double* lin /* lc0, ldc[3], llap, luo[3], lbeta, lkappa, lmob */) {
    const Phys<double> ph = unpack(phys);
    double dcc[3] = {dc[0], dc[1], L::D == 3 ? dc[2] : 0.0};
    double uoo[3] = {uo[0], uo[1], L::D == 3 ? uo[2] : 0.0};
    double luu[3] = {lu[0], lu[1], L::D == 3 ? lu[2] : 0.0};
    NodeAdj<L, double> o{};
    node_reverse<L, double>(ph, fp, gp, c0, dcc, lap, uoo, lfs, lgs, luu, o);
    for (int q = 0; q < L::Q; ++q) {
        lfp[q] = o.lfp[q];
        lgp[q] = o.lgp[q];
    }
    lin[0] = o.lc0;
    for (int d = 0; d < 3; ++d) lin[1 + d] = o.ldc[d];
    lin[4] = o.llap;
    for (int d = 0; d < 3; ++d) lin[5 + d] = o.luo[d];
    lin[8] = o.lbeta;
    lin[9] = o.lkappa;
    lin[10] = o.lmob;
}

// Initial state + first collision at one node (u = 0, initial p).
template <class L>
void init_fwd(const double* phys, double sigma, double half, double c0, const double* dc, double lap, double* fs,
              double* gs) {
    const Phys<double> ph = unpack(phys);
    const NodeScalars<double> z = node_scalars<L>(ph, c0, lap);
    const double k = ph.case_type == CLIP_CASE_DROP ? 1.0 : (ph.case_type == CLIP_CASE_BUBBLE ? -1.0 : 0.0);
    const double p = k * c0 * sigma / half / (z.rho / 3.0);
    double dcc[3] = {dc[0], dc[1], L::D == 3 ? dc[2] : 0.0};
    double nr[3], rt, mag;
    normal_of<L>(dcc, nr, rt, mag);
    double fp[L::Q], gp[L::Q];
    const double gi = 2.0 * c0 * (1.0 - c0) / ph.W;
    for (int q = 0; q < L::Q; ++q) {
        fp[q] = p * L::w(q);
        gp[q] = L::w(q) * (c0 - gi * edot<L>(q, nr));
    }
    double mom[3], Pf[3][3];
    pop_moments<L>(fp, mom, Pf);
    const double u0[3] = {0, 0, 0};
    collide_node<L, double>(
        z, p, Pf, c0, dcc, u0, gp, [&](int q, double v) { gs[q] = v; }, [&](int q, double v) { fs[q] = v; });
}

template <class L>
void init_rev(const double* phys, double sigma, double half, double c0, const double* dc, double lap,
              const double* lfs, const double* lgs, double* lin /* lc0, ldc[3], llap, lbeta, lkappa, lmob, lsigma */) {
    const Phys<double> ph = unpack(phys);
    double dcc[3] = {dc[0], dc[1], L::D == 3 ? dc[2] : 0.0};
    double lc0 = 0, ldc[3] = {0, 0, 0}, llap = 0, lb = 0, lk = 0, lm = 0, ls = 0;
    init_reverse<L, double>(ph, sigma, half, c0, dcc, lap, lfs, lgs, lc0, ldc, llap, lb, lk, lm, ls);
    lin[0] = lc0;
    for (int d = 0; d < 3; ++d) lin[1 + d] = ldc[d];
    lin[4] = llap;
    lin[5] = lb;
    lin[6] = lk;
    lin[7] = lm;
    lin[8] = ls;
}

}  // namespace

extern "C" {
int node_fwd(int dim, const double* phys, const double* fp, const double* gp, double c0, const double* dc, double lap,
             const double* uo, double* fs, double* gs, double* u) {
    if (dim == 2) fwd<D2Q9>(phys, fp, gp, c0, dc, lap, uo, fs, gs, u);
    else fwd<D3Q19>(phys, fp, gp, c0, dc, lap, uo, fs, gs, u);
    return 0;
}
int node_rev(int dim, const double* phys, const double* fp, const double* gp, double c0, const double* dc,
             double lap, const double* uo, const double* lfs, const double* lgs, const double* lu, double* lfp,
             double* lgp, double* lin) {
    if (dim == 2) rev<D2Q9>(phys, fp, gp, c0, dc, lap, uo, lfs, lgs, lu, lfp, lgp, lin);
    else rev<D3Q19>(phys, fp, gp, c0, dc, lap, uo, lfs, lgs, lu, lfp, lgp, lin);
    return 0;
}
int node_rev32(int dim, const double* phys, const double* fp, const double* gp, double c0, const double* dc,
               double lap, const double* uo, const double* lfs, const double* lgs, const double* lu, double nfloor,
               double* lfp, double* lgp, double* lin) {
    if (dim == 2) rev32<D2Q9>(phys, fp, gp, c0, dc, lap, uo, lfs, lgs, lu, nfloor, lfp, lgp, lin);
    else rev32<D3Q19>(phys, fp, gp, c0, dc, lap, uo, lfs, lgs, lu, nfloor, lfp, lgp, lin);
    return 0;
}
// x[i] / d through div_rn with r = 1/d (for the division tests).
int div_rn_batch(const double* x, long long n, double d, double* out) {
    const double r = 1.0 / d;
    for (long long i = 0; i < n; ++i) out[i] = div_rn(x[i], d, r);
    return 0;
}
int div_rn_pairs(const double* x, const double* d, long long n, double* out) {
    for (long long i = 0; i < n; ++i) out[i] = div_rn(x[i], d[i], 1.0 / d[i]);
    return 0;
}
int init_fwd(int dim, const double* phys, double sigma, double half, double c0, const double* dc, double lap,
             double* fs, double* gs) {
    if (dim == 2) init_fwd<D2Q9>(phys, sigma, half, c0, dc, lap, fs, gs);
    else init_fwd<D3Q19>(phys, sigma, half, c0, dc, lap, fs, gs);
    return 0;
}
int init_rev(int dim, const double* phys, double sigma, double half, double c0, const double* dc, double lap,
             const double* lfs, const double* lgs, double* lin) {
    if (dim == 2) init_rev<D2Q9>(phys, sigma, half, c0, dc, lap, lfs, lgs, lin);
    else init_rev<D3Q19>(phys, sigma, half, c0, dc, lap, lfs, lgs, lin);
    return 0;
}
}

// ---------------------------------------------------------------- adjoint maps
// The host-built transposes of the reverse step (adjoint_maps.hpp) on a dense
// grid, for the CPU tests against the oracle's boundary phases.
#include "adjoint_maps.hpp"

namespace {
template <class L>
int pull_eval(const clip_config* cfg, const double* F, const double* prev, const double* u, double* out) {
    const DevGrid g = grid_of_config(*cfg, L::D, 1);
    const DevBC bc = bc_of_config(*cfg, L::D);
    const auto vf = inlet_flags(*cfg, g, L::D);
    long long poff[6], tot = 0;
    for (int s = 0; s < 6; ++s) {
        poff[s] = tot;
        if (s < 2 * L::D && cfg->bc_type[s] == CLIP_BC_FREE_CONVECT) tot += plane_size(g, s) * L::Q;
    }
    SymPuller<L> sp{g, bc, vf};
    std::vector<PullTerm> terms;
    for (long long i = 0; i < g.nodes * L::Q; ++i) out[i] = 0.0;
    for (int k = g.dmin[2]; k <= g.dmax[2]; ++k)
        for (int j = g.dmin[1]; j <= g.dmax[1]; ++j)
            for (int i = g.dmin[0]; i <= g.dmax[0]; ++i) {
                const int c[3] = {i, j, k};
                bool shell = false;
                for (int a = 0; a < L::D; ++a) shell |= !bc.both_periodic[a] && (c[a] == g.dmin[a] || c[a] == g.dmax[a]);
                const long long x = g.idx(i, j, k);
                for (int q = 0; q < L::Q; ++q) {
                    double v = 0.0;
                    if (!shell) {
                        int s3[3] = {i - L::ex(q), j - L::ey(q), k - L::ez(q)};
                        for (int a = 0; a < L::D; ++a) {
                            const int span = g.dmax[a] - g.dmin[a] + 1;
                            if (s3[a] < g.dmin[a]) s3[a] += span;
                            if (s3[a] > g.dmax[a]) s3[a] -= span;
                        }
                        v = F[(long long)q * g.nodes + g.idx(s3[0], s3[1], s3[2])];
                    } else {
                        terms.clear();
                        sp.lvl3(c, q, terms);
                        for (const auto& t : terms) {
                            double w = 1.0;
                            if (t.fc >= 0) {
                                const double un = std::fabs(u[(long long)(t.fc / 2) * g.nodes + t.nb]);
                                w = (t.branch == 0 ? un : 1.0) / (1.0 + un);
                            }
                            v += w * (t.kind == 1 ? prev[poff[t.side] + t.idx] : F[t.idx]);
                        }
                    }
                    out[(long long)q * g.nodes + x] = v;
                }
            }
    return 0;
}
}  // namespace

extern "C" {
// f_post at every interior node from the stored populations F (SoA, dense,
// ghosts included, inlet ghost values in place), the outflow planes (free-
// convect sides concatenated, q * plane + t each) and u (SoA): the bulk pull
// and the boundary chain's terms.
int maps_pull_eval(const clip_config* cfg, const double* F, const double* prev, const double* u, double* out) {
    return cfg->dim == 2 ? pull_eval<D2Q9>(cfg, F, prev, u, out) : pull_eval<D3Q19>(cfg, F, prev, u, out);
}
// Source of every node of the new phi under one ghost refresh (-2: not written).
int maps_ring(const clip_config* cfg, long long* src) {
    const int D = cfg->dim;
    const DevGrid g = grid_of_config(*cfg, D, 1);
    const DevBC bc = bc_of_config(*cfg, D);
    const auto vf = inlet_flags(*cfg, g, D);
    const std::vector<long long> s = ring_sources(g, bc, D, vf);
    for (long long i = 0; i < g.nodes; ++i) src[i] = s[i];
    return 0;
}
}
