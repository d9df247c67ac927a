// adjoint_maps.hpp — host-built sparse transposes for the reverse step on
// domains with non-periodic sides.
//
// The forward pull of a boundary-layer node resolves the reference's
// streaming + boundary chain (Puller, kernels.hpp: inlet, half bounce-back,
// free convect, Neumann).  Each (node, direction) reads at most a few stored
// values: a population of the state, an outflow-plane entry (the previous
// step's f_post, kernelFreeConvect Solver.cu:342-453), or an inlet ghost
// value (JetBoundary, Solver.cu:550-921, constant in time, a function of the
// inlet velocity).  SymPuller evaluates the same chain symbolically and the
// lists are inverted into CSR form (one row per target), so that the reverse
// kernels gather every target's adjoint in a fixed order.
//
// The ghost ring of the new phi (k_ghost_resolve / k_ghost_velocity) is a
// copy map from the previous phi; ring_sources simulates the three passes
// with labels to get, for every ring cell, the previous-phi cell it copies
// (interior or ghost) or none (never rewritten: constant).
#pragma once

#include "geometry.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <map>
#include <vector>

namespace clipamd {

// Host copy of the inlet disc flags: per velocity side, 1 for a ghost-plane
// cell inside geometry 0 (JetBoundary's sdf <= 0 test).
inline std::vector<std::vector<unsigned char>> inlet_flags(const clip_config& cfg, const DevGrid& g, int D) {
    std::vector<std::vector<unsigned char>> vf(6);
    for (int s = 0; s < 2 * D; ++s) {
        if (cfg.bc_type[s] != CLIP_BC_VELOCITY) continue;
        const long long ps = plane_size(g, s);
        vf[s].resize(ps);
        const int a = s / 2;
        const int t1 = (a == 0) ? 1 : 0, t2 = (a == 2) ? 1 : 2;
        for (long long t = 0; t < ps; ++t) {
            int cc[3];
            cc[t1] = (int)(t % g.ext[t1]);
            cc[t2] = (int)(t / g.ext[t1]);
            cc[a] = (s & 1) ? g.dmax[a] + 1 : g.dmin[a] - 1;
            const double z = (D == 3) ? (double)cc[2] : 0.0;
            vf[s][t] = sdf(cfg, 0, cc[0], cc[1], z) <= 0 ? 1 : 0;
        }
    }
    return vf;
}

// One stored value a pulled population depends on.
struct PullTerm {
    int kind;       // 0: state population q*nodes + n; 1: outflow-plane entry; 2: inlet ghost value q*nodes + n;
                    // 3: any other ghost value (a constant of the initial state, no adjoint)
    long long idx;  // kind 0 / 2: population index; kind 1: entry of side `side`'s plane set (q * plane + t)
    int side;       // kind 1: the free-convect side
    int fc;         // -1: weight 1; else the free-convect side whose |u_n(nb)| weights the term
    long long nb;   // free-convect: device node of the inner neighbour
    int branch;     // free-convect: 0 = |u| / (1 + |u|) (the pulled value), 1 = 1 / (1 + |u|) (the plane value)
};

// Host mirror of Puller's chain (same branch order and conditions).
template <class L>
struct SymPuller {
    const DevGrid& g;
    const DevBC& bc;
    const std::vector<std::vector<unsigned char>>& vflag;  // per side, host copy of the inlet disc flags

    bool interior(const int s[3]) const { return g.interior(s, L::D); }

    void load(int q, const int s[3], int fc, long long nb, int branch, std::vector<PullTerm>& out) const {
        const long long node = g.idx(s[0], s[1], s[2]);
        if (interior(s)) {
            out.push_back({0, (long long)q * g.nodes + node, -1, fc, nb, branch});
            return;
        }
        // a ghost cell: an inlet value inside a disc, otherwise a constant of the
        // initial state
        int kind = 3;
        for (int side = 2 * L::D - 1; side >= 0; --side) {
            const int a = side >> 1;
            if (bc.type[side] != BC_VELOCITY) continue;
            const int ghost = (side & 1) ? g.dmax[a] + 1 : g.dmin[a] - 1;
            if (s[a] != ghost) continue;
            if (vflag[side][plane_index(g, side, s)]) kind = 2;
            break;
        }
        out.push_back({kind, (long long)q * g.nodes + node, -1, fc, nb, branch});
    }

    void raw(const int c[3], int q, int fc, long long nb, int branch, std::vector<PullTerm>& out) const {
        int s[3] = {c[0] - L::ex(q), c[1] - L::ey(q), c[2] - L::ez(q)};
        for (int a = 0; a < L::D; ++a)
            if (bc.periodic[a]) {
                if (s[a] < g.dmin[a]) s[a] = g.dmax[a];
                else if (s[a] > g.dmax[a]) s[a] = g.dmin[a];
            }
        if (bc.is_velocity) {
            for (int side = 2 * L::D - 1; side >= 0; --side) {
                const int a = side >> 1;
                const int inward = (side & 1) ? -1 : 1;
                if (bc.type[side] != BC_VELOCITY || L::e(q, a) != inward) continue;
                const int ghost = (side & 1) ? g.dmax[a] + 1 : g.dmin[a] - 1;
                if (s[a] != ghost) continue;
                if (!vflag[side][plane_index(g, side, s)]) {
                    int b[3] = {s[0], s[1], s[2]};
                    b[a] = (side & 1) ? g.dmax[a] : g.dmin[a];
                    load(L::opp(q), b, fc, nb, branch, out);
                    return;
                }
                break;
            }
        }
        load(q, s, fc, nb, branch, out);
    }

    void lvl1(const int c[3], int q, int fc, long long nb, int branch, std::vector<PullTerm>& out) const {
        const int src = wall_source<L>(g, bc, c, q);
        if (src >= 0) {
            load(src, c, fc, nb, branch, out);
            return;
        }
        raw(c, q, fc, nb, branch, out);
    }

    void lvl2(const int c[3], int q, std::vector<PullTerm>& out) const {
        if (bc.is_fc) {
            for (int side = 2 * L::D - 1; side >= 0; --side) {
                const int a = side >> 1;
                const int inward = (side & 1) ? -1 : 1;
                if (bc.type[side] != BC_FREE_CONVECT || L::e(q, a) != inward ||
                    c[a] != ((side & 1) ? g.dmax[a] : g.dmin[a]))
                    continue;
                int nbc[3] = {c[0], c[1], c[2]};
                nbc[a] += inward;
                const long long nb = g.idx(nbc[0], nbc[1], nbc[2]);
                out.push_back({1, (long long)q * plane_size(g, side) + plane_index(g, side, c), side, side, nb, 1});
                lvl1(nbc, q, side, nb, 0, out);
                return;
            }
        }
        lvl1(c, q, -1, -1, 0, out);
    }

    void lvl3(const int c[3], int q, std::vector<PullTerm>& out) const {
        if (bc.is_neumann) {
            for (int side = 2 * L::D - 1; side >= 0; --side) {
                const int a = side >> 1;
                const int inward = (side & 1) ? -1 : 1;
                if (bc.type[side] != BC_NEUMANN || L::e(q, a) != inward ||
                    c[a] != ((side & 1) ? g.dmax[a] : g.dmin[a]))
                    continue;
                int nbc[3] = {c[0], c[1], c[2]};
                nbc[a] += inward;
                lvl2(nbc, q, out);
                return;
            }
        }
        lvl2(c, q, out);
    }
};

// One reader's contribution to a target.  rd: the reader's adjoint index
// (q*nodes + x).  Pull rows: uidx = a*nodes + nb of a free-convect weight
// (-1: weight 1) and branch its kind (see PullTerm).  Velocity rows (the
// derivative of a free-convect pull a = (P + |u| X) / (1 + |u|) with respect
// to u_n(nb), sign(u) (X - P) / (1 + |u|)^2): xidx = population index of X in
// the state, side / pidx = the outflow-plane entry P.
struct AdjEntry {
    long long rd;
    long long uidx;
    long long xidx;
    long long pidx;
    int branch;
    int side;
};

struct Csr {
    std::vector<long long> target;  // target index per row
    std::vector<int> start;         // row r covers entries [start[r], start[r+1])
    std::vector<AdjEntry> e;
};

inline Csr make_csr(std::map<long long, std::vector<AdjEntry>>& rows) {
    Csr c;
    c.start.push_back(0);
    for (auto& kv : rows) {
        c.target.push_back(kv.first);
        for (const auto& x : kv.second) c.e.push_back(x);
        c.start.push_back((int)c.e.size());
    }
    return c;
}

// Ring map of one ghost refresh: for every node of the new phi that the
// refresh writes, the node of the old phi it copies (the passes of
// SolverImpl::ghost_refresh in order; a read from the new phi of a cell
// written by an earlier pass resolves to that cell's source).
inline std::vector<long long> ring_sources(const DevGrid& g, const DevBC& bc, int D,
                                           const std::vector<std::vector<unsigned char>>& vflag) {
    std::vector<long long> src((size_t)g.nodes, -2);  // -2: not written (keeps its constant)
    auto each_ghost = [&](auto fn) {
        for (int side = 0; side < 2 * D; ++side) {
            const int a = side >> 1;
            const int t1 = (a == 0) ? 1 : 0, t2 = (a == 2) ? 1 : 2;
            for (long long t = 0; t < plane_size(g, side); ++t) {
                int c[3];
                c[t1] = (int)(t % g.ext[t1]);
                c[t2] = (int)(t / g.ext[t1]);
                c[a] = (side & 1) ? g.dmax[a] + 1 : g.dmin[a] - 1;
                fn(side, t, c);
            }
        }
    };
    auto resolve = [&](int do_periodic, int do_mirror, int from_new, std::vector<long long>& cur) {
        std::vector<long long> nxt = cur;
        each_ghost([&](int, long long, const int c[3]) {
            int s[3] = {c[0], c[1], c[2]};
            bool changed = false;
            for (int a = 0; a < D; ++a) {
                const bool lo = c[a] < g.dmin[a], hi = c[a] > g.dmax[a];
                if (!lo && !hi) continue;
                const int ty = bc.type[2 * a + (hi ? 1 : 0)];
                const bool mirror = bc.mirror_gate && (ty == BC_WALL || ty == BC_FREE_CONVECT || ty == BC_NEUMANN ||
                                                       ty == BC_SLIP_WALL || ty == BC_DO_NOTHING);
                if (do_mirror && mirror) {
                    s[a] = lo ? g.dmin[a] : g.dmax[a];
                    changed = true;
                } else if (do_periodic && bc.periodic[a] && !(do_mirror && mirror)) {
                    s[a] = lo ? g.dmax[a] : g.dmin[a];
                    changed = true;
                }
            }
            if (!changed) return;
            const long long si = g.idx(s[0], s[1], s[2]);
            nxt[g.idx(c[0], c[1], c[2])] = (!g.interior(s, D) && from_new) ? cur[si] : si;
        });
        cur = nxt;
    };
    if (!bc.is_velocity) {
        resolve(1, 1, 0, src);
        return src;
    }
    resolve(1, 0, 0, src);
    std::vector<long long> nxt = src;
    each_ghost([&](int side, long long t, const int c[3]) {
        if (bc.type[side] != BC_VELOCITY || vflag[side][t]) return;
        const int a = side >> 1;
        int b[3] = {c[0], c[1], c[2]};
        b[a] = (side & 1) ? g.dmax[a] : g.dmin[a];
        const long long bi = g.idx(b[0], b[1], b[2]);
        // a boundary cell that is itself a ghost (edge of the plane) is read
        // from the new phi (k_ghost_velocity's ghost_source)
        nxt[g.idx(c[0], c[1], c[2])] = g.interior(b, D) ? bi : src[bi];
    });
    src = nxt;
    if (bc.mirror_gate) resolve(0, 1, 1, src);
    return src;
}

}  // namespace clipamd
