// launchers.hpp — the reference Solver's boundary launchers
// (include/solver/Solver.cuh:61-107, src/solver/Solver.cu:985-1062) as HIP
// kernels on caller device arrays in CLIP's own layout: FP64, node-major AoS
// ((i * ext_y + j) * ext_z + k) * ndof + dof with the ghost ring included
// (Domain.cuh:63-68).
//
// The solver's step never calls these: it resolves every boundary inside the
// pull of its fused step kernels (kernels.hpp).  They exist for code written
// against CLIP's Solver class that applies a boundary to arrays of its own.
// Where the reference's kernel reads values that the same launch writes
// (periodic corners; the inner neighbours of free-convect and Neumann nodes
// on two sides), these read a copy taken before the launch, which is the
// result the reference gets when no such overlap occurs.
#pragma once

#include "geometry.hpp"
#include "kernels.hpp"

namespace clipamd {

// CLIP's dense extents (no row padding) and node index.
struct AosGrid {
    int ext[3], dmin[3], dmax[3];
    int D;
    __host__ __device__ long long node(int i, int j, int k) const {
        return ((long long)i * ext[1] + j) * ext[2] + k;
    }
    __host__ __device__ long long count() const { return (long long)ext[0] * ext[1] * ext[2]; }
    __host__ __device__ void coords(long long t, int c[3]) const {
        c[2] = (int)(t % ext[2]);
        t /= ext[2];
        c[1] = (int)(t % ext[1]);
        c[0] = (int)(t / ext[1]);
    }
    __host__ __device__ bool interior(const int c[3]) const {
        bool in = true;
        for (int a = 0; a < 3; ++a) in = in && c[a] >= dmin[a] && c[a] <= dmax[a];
        return in;
    }
};

inline AosGrid aos_grid(const clip_config& cfg) {
    const int D = cfg.dim;
    const DevGrid g = grid_of_config(cfg, D, 1);
    AosGrid a{};
    for (int k = 0; k < 3; ++k) {
        a.ext[k] = g.ext[k];
        a.dmin[k] = g.dmin[k];
        a.dmax[k] = g.dmax[k];
    }
    a.D = D;
    return a;
}

// One direction list per side (X-, X+, Y-, Y+, Z-, Z+) of a reference map.
struct SideLists {
    int dir[6][5];
    int n;
};
template <class L>
inline SideLists side_lists(bool slip) {
    SideLists s{};
    s.n = L::NIN;
    for (int side = 0; side < 2 * L::D; ++side)
        for (int q = 0; q < L::Q; ++q) {
            const int m = slip ? L::slip_pos(side, q) : L::wall_pos(side, q);
            if (m >= 0) s.dir[side][m] = q;
        }
    return s;
}

// kernelPeriodicBoundary (Solver.cu:75-141) for one axis: over the plane
// normal to `ax` (every other coordinate, ghosts included) the two ghost
// cells take the opposite interior layer.  Launched for X, then Y, then Z:
// the full periodic wrap, which the reference's single launch leaves racy
// at edges and corners.
__global__ void k_aos_periodic(double* a, int ndof, AosGrid g, int ax) {
    const int t1 = (ax == 0) ? 1 : 0, t2 = (ax == 2) ? 1 : 2;
    const long long cells = (long long)g.ext[t1] * g.ext[t2];
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cells * ndof) return;
    const int d = (int)(t % ndof);
    t /= ndof;
    int c[3];
    c[t1] = (int)(t % g.ext[t1]);
    c[t2] = (int)(t / g.ext[t1]);
    int lo[3] = {c[0], c[1], c[2]}, hi[3] = {c[0], c[1], c[2]}, slo[3] = {c[0], c[1], c[2]},
        shi[3] = {c[0], c[1], c[2]};
    lo[ax] = g.dmin[ax] - 1;
    hi[ax] = g.dmax[ax] + 1;
    slo[ax] = g.dmax[ax];
    shi[ax] = g.dmin[ax];
    a[g.node(lo[0], lo[1], lo[2]) * ndof + d] = a[g.node(slo[0], slo[1], slo[2]) * ndof + d];
    a[g.node(hi[0], hi[1], hi[2]) * ndof + d] = a[g.node(shi[0], shi[1], shi[2]) * ndof + d];
}

__device__ inline bool on_side(const AosGrid& g, int side, const int c[3]) {
    const int a = side >> 1;
    return c[a] == ((side & 1) ? g.dmax[a] : g.dmin[a]);
}

// kernelHalfBounceBack (Solver.cu:248-333) with the map `lists` (wall or
// slip): at every interior node, for list position m, for each side of type
// Wall or SlipWall the node touches (X- first), a_post[list[s][m]] =
// a[list[s ^ 1][m]] at the same node; the same for b when given.
__global__ void k_aos_half_bb(const double* a, double* a_post, const double* b, double* b_post, AosGrid g,
                              DevBC bc, SideLists lists, int Q) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.count()) return;
    int c[3];
    g.coords(t, c);
    if (!g.interior(c)) return;
    for (int m = 0; m < lists.n; ++m)
        for (int s = 0; s < 2 * g.D; ++s) {
            if (!(bc.type[s] == BC_WALL || bc.type[s] == BC_SLIP_WALL) || !on_side(g, s, c)) continue;
            const long long dst = t * Q + lists.dir[s][m], src = t * Q + lists.dir[s ^ 1][m];
            a_post[dst] = a[src];
            if (b) b_post[dst] = b[src];
        }
}

// kernelFreeConvect (Solver.cu:342-453) at the interior nodes of
// free-convect sides: a[x, q] = (a_prev[x, q] + |u_n| a[x + n, q]) / (1 + |u_n|)
// for the side's in-directions q (wallBCMap order, sides X- first), n the
// inward normal and u_n the normal velocity at x + n.  a[x + n] is read from
// a_in, a copy of a taken before the launch.
__global__ void k_aos_free_convect(const double* vel, double* a, const double* a_in, const double* a_prev, double* b,
                                   const double* b_in, const double* b_prev, AosGrid g, DevBC bc, SideLists lists,
                                   int Q) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.count()) return;
    int c[3];
    g.coords(t, c);
    if (!g.interior(c)) return;
    for (int s = 0; s < 2 * g.D; ++s) {
        if (bc.type[s] != BC_FREE_CONVECT || !on_side(g, s, c)) continue;
        const int ax = s >> 1;
        int nb[3] = {c[0], c[1], c[2]};
        nb[ax] += (s & 1) ? -1 : 1;
        const long long ni = g.node(nb[0], nb[1], nb[2]);
        const double cv = fabs(vel[ni * g.D + ax]);
        for (int m = 0; m < lists.n; ++m) {
            const int q = lists.dir[s][m];
            a[t * Q + q] = (a_prev[t * Q + q] + cv * a_in[ni * Q + q]) / (1.0 + cv);
            if (b) b[t * Q + q] = (b_prev[t * Q + q] + cv * b_in[ni * Q + q]) / (1.0 + cv);
        }
    }
}

// kernelNeumann (Solver.cu:462-542): a[x, q] = a[x + n, q] for the
// in-directions of Neumann sides, read from the copy a_in.
__global__ void k_aos_neumann(double* a, const double* a_in, double* b, const double* b_in, AosGrid g, DevBC bc,
                              SideLists lists, int Q) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.count()) return;
    int c[3];
    g.coords(t, c);
    if (!g.interior(c)) return;
    for (int s = 0; s < 2 * g.D; ++s) {
        if (bc.type[s] != BC_NEUMANN || !on_side(g, s, c)) continue;
        const int ax = s >> 1;
        int nb[3] = {c[0], c[1], c[2]};
        nb[ax] += (s & 1) ? -1 : 1;
        const long long ni = g.node(nb[0], nb[1], nb[2]);
        for (int m = 0; m < lists.n; ++m) {
            const int q = lists.dir[s][m];
            a[t * Q + q] = a_in[ni * Q + q];
            if (b) b[t * Q + q] = b_in[ni * Q + q];
        }
    }
}

// JetBoundary (Solver.cu:550-921) on the ghost plane of velocity side
// `side`.  Inside the inlet disc (flag set): the in-directions a of f and g
// from the outgoing and tangential ghost populations,
//   f_a = f_abar + Gamma_a(U) - Gamma_abar(U) - (1 / weight) e_t,a . M_t,
//   g_a = geq_a - (g_abar - geq_abar) - [a not normal] N / weight,
// with U the side's inlet velocity at the cell (Uplane), geq = phi (Gamma +
// w), M_t = sum_{q >= 1, e_n = 0} e_t (f_q - Gamma_q) and N = sum_{q >= 1,
// e_n = 0} (g_q - geq_q).  Outside: the boundary node's outgoing populations
// reversed and its phi.
template <class L>
__global__ void k_aos_velocity(double* c, double* f, double* gg, AosGrid g, int side, const unsigned char* flag,
                               const double* Uplane, SideLists lists) {
    constexpr int Q = L::Q;
    const int ax = side >> 1;
    const int t1 = (ax == 0) ? 1 : 0, t2 = (ax == 2) ? 1 : 2;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.ext[t1] * g.ext[t2]) return;
    int gc[3];
    gc[t1] = (int)(t % g.ext[t1]);
    gc[t2] = (int)(t / g.ext[t1]);
    gc[ax] = (side & 1) ? g.dmax[ax] + 1 : g.dmin[ax] - 1;
    const long long gi = g.node(gc[0], gc[1], gc[2]);
    // plane cell index of the solver's flag / velocity planes (t1 fastest)
    const long long pt = (long long)gc[t2] * g.ext[t1] + gc[t1];
    if (flag[pt]) {
        const double U[3] = {Uplane[3 * pt], Uplane[3 * pt + 1], Uplane[3 * pt + 2]};
        const double cg = c[gi];
        double feq[Q], geq[Q], Mt[3] = {0, 0, 0}, N = 0;
        for (int q = 1; q < Q; ++q) {
            feq[q] = gamma_u<L, double>(q, U);
            geq[q] = cg * (feq[q] + L::w(q));
            if (L::e(q, ax) == 0) {
                for (int d = 0; d < L::D; ++d)
                    if (d != ax) Mt[d] += L::e(q, d) * (f[gi * Q + q] - feq[q]);
                N += gg[gi * Q + q] - geq[q];
            }
        }
        const double weight = L::NIN - 1;
        for (int m = 0; m < lists.n; ++m) {
            const int a = lists.dir[side][m], b = L::opp(a);
            double corr = 0;
            for (int d = 0; d < L::D; ++d)
                if (d != ax) corr += L::e(a, d) * Mt[d];
            f[gi * Q + a] = f[gi * Q + b] + feq[a] - feq[b] - (1.0 / weight) * corr;
            double gv = geq[a] - (gg[gi * Q + b] - geq[b]);
            if (m != 0) gv -= N / weight;
            gg[gi * Q + a] = gv;
        }
    } else {
        int bn[3] = {gc[0], gc[1], gc[2]};
        bn[ax] = (side & 1) ? g.dmax[ax] : g.dmin[ax];
        const long long bi = g.node(bn[0], bn[1], bn[2]);
        for (int m = 0; m < lists.n; ++m) {
            const int a = lists.dir[side][m];
            gg[gi * Q + a] = gg[bi * Q + L::opp(a)];
            f[gi * Q + a] = f[bi * Q + L::opp(a)];
        }
        c[gi] = c[bi];
    }
}

// kernelMirrorBoundary (Solver.cu:927-978) for one side: every cell of the
// side's ghost plane takes the value of the boundary cell next to it.
__global__ void k_aos_mirror(double* c, AosGrid g, int side) {
    const int ax = side >> 1;
    const int t1 = (ax == 0) ? 1 : 0, t2 = (ax == 2) ? 1 : 2;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.ext[t1] * g.ext[t2]) return;
    int gc[3];
    gc[t1] = (int)(t % g.ext[t1]);
    gc[t2] = (int)(t / g.ext[t1]);
    gc[ax] = (side & 1) ? g.dmax[ax] + 1 : g.dmin[ax] - 1;
    int bn[3] = {gc[0], gc[1], gc[2]};
    bn[ax] = (side & 1) ? g.dmax[ax] : g.dmin[ax];
    c[g.node(gc[0], gc[1], gc[2])] = c[g.node(bn[0], bn[1], bn[2])];
}

// Streaming copy, 16 bytes per lane and one element per thread: the device's
// practical HBM rate for a pure read + write stream (clip_copy_rate).
struct alignas(16) Copy16 {
    unsigned w[4];
};
__global__ void k_copy16(const Copy16* __restrict__ a, Copy16* __restrict__ b, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

}  // namespace clipamd
