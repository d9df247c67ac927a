// adjoint.hpp — reverse-mode (adjoint) kernels of the coupled step, FP64 or
// FP32 populations stored plainly, periodic or bounded by walls, velocity
// inlets, free-convect outflows, Neumann and do-nothing sides.
//
// Forward step n (state S^n = f*, g*, u, phi with its ghost ring, outflow
// planes P = the previous step's f_post / g_post on free-convect sides):
//   f_post/g_post(x) = f*/g*(x - e_q) in the bulk; the boundary chain
//                    (Puller: inlet, bounce-back, free convect, Neumann) in the
//                    boundary layer, reading S^n's populations, planes and u
//   phi'(x)          = sum_q g_post(x) at interior nodes; the ghost ring of
//                    phi' copies cells of phi^n (k_ghost_*: the lagged ring)
//   stencil(x)       reads phi' at the 27 neighbours (ring cells included)
//   node(x)          u, g*, f* from f_post, g_post, phi', grad, lap, u^n
//   P^{n+1}          f_post / g_post on the outflow boundary nodes
// The reverse sweep of one step runs:
//   k_adj_node   node VJP (node.hpp) at every interior node (+ the adjoint
//                of P^{n+1} on outflow boundary nodes); a bulk node scatters
//                its f_post adjoint through the transpose of its pull (a
//                permutation) straight into the adjoint of S^n's f*
//   k_adj_phi    transpose of the stencil and of the phase sum's consumer:
//                adjoint of phi' at every cell (interior and ring); a bulk
//                node then scatters its g_post adjoint plus its phi' adjoint
//                into the adjoint of S^n's g*
//   k_adj_ring   transpose of the ring copy: adjoint of phi^n (host-built map)
//   k_adj_csr_*  transposes of the boundary-layer chains (host-built lists):
//                populations, outflow planes, inlet values, and u through the
//                free-convect weights
// Every entry of the adjoint of S^n's populations has exactly one reader in
// the step: a bulk node (its scatter writes the entry), a boundary-layer node
// (it zeroes the entry, the row lists then add the chain's terms) or none
// (the entry's own boundary-layer node zeroes it).
#pragma once

#include "adjoint_maps.hpp"
#include "kernels.hpp"
#include "node.hpp"

namespace clipamd {

// Weights of the phi stencil (phi_stencil): d(grad_a)/d phi(x+d) and
// d(lap)/d phi(x+d) for offset d in {-1,0,1}^D.
template <int D, class T>
__device__ __forceinline__ void stencil_weights(int dx, int dy, int dz, T wg[3], T& wl) {
    if (D == 2) {
        const int m = (dx != 0) + (dy != 0);
        const T g1 = (m == 1) ? T(1.0 / 3.0) : T(1.0 / 12.0);
        wg[0] = dx ? T(dx) * g1 : T(0);
        wg[1] = dy ? T(dy) * g1 : T(0);
        wg[2] = T(0);
        wl = (m == 0) ? T(-20.0 / 6.0) : (m == 1 ? T(4.0 / 6.0) : T(1.0 / 6.0));
    } else {
        const int m = (dx != 0) + (dy != 0) + (dz != 0);
        const int off[3] = {dx, dy, dz};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int others = m - (off[a] != 0);
            const T c = others == 0 ? T(4.0 / 9.0) : (others == 1 ? T(1.0 / 9.0) : T(1.0 / 36.0));
            wg[a] = off[a] ? T(0.5) * T(off[a]) * c : T(0);
        }
        wl = (m == 0) ? T(-200.0 / 48.0) : (m == 1 ? T(20.0 / 48.0) : (m == 2 ? T(6.0 / 48.0) : T(1.0 / 48.0)));
    }
}

// Row lists (CSR) of a host-built transpose, in device memory.
struct AdjRows {
    const long long* target;
    const int* start;
    const AdjEntry* e;
    int rows;
};

// Adjoint of grad phi (dc[2] = 0 in 2D) and of lap phi at one node, one
// 16-byte (FP32) or 32-byte (FP64) record: the stencil transpose reads a
// neighbour's record with wide loads instead of D + 1 separate planes.
template <class CT>
struct alignas(4 * sizeof(CT)) StAdj {
    CT dc[3];
    CT lap;
};

template <class CT>
struct AdjArgs {
    DevGrid g;
    DevBC bc;
    Phys<CT> ph;
    const CT* F;     // f* of S^n (StorePlain), Q x nodes
    const CT* G;     // g* of S^n
    const CT* phin;  // phi^{n+1} with the ghost ring of phi^n (or phi0 at init)
    const CT* uo;    // u^n
    const CT* lF;    // adjoint of f*^{n+1}
    const CT* lG;    // adjoint of g*^{n+1}
    CT* lFo;         // adjoint of f*^n (written through the pull's transpose)
    CT* lGo;         // adjoint of g*^n
    CT* lU;          // in: adjoint of u^{n+1};  out: adjoint of u^n
    CT* lPhi;        // in: adjoint of phi^{n+1}; out: adjoint of phi^n (interior)
    CT* lfp;         // scratch Q x nodes: adjoint of f_post
    CT* lgp;         // scratch Q x nodes: adjoint of g_post
    CT* lst;         // scratch, nodes: adjoint of phi at the node (the stencil's centre value)
    StAdj<CT>* lsv;  // scratch, nodes: adjoint of grad phi and lap phi at the node
    CT* lphin;       // scratch nodes: adjoint of phi^{n+1}
    double* part;    // 4 x blocks partial sums: beta, kappa, mobility, sigma
    CT sigma, half_length;
    CT nfloor;       // regularisation of the normal's derivative (0 = exact)
    // outflow planes (free-convect sides): S^n's planes, the adjoint of
    // P^{n+1} (in) and of P^n (out); each a concatenation of the sides' plane
    // sets at offsets poff[side]
    const CT* pf;
    const CT* pg;
    const CT* lPf_in;
    const CT* lPg_in;
    CT* lPf_out;
    CT* lPg_out;
    long long poff[6];
};

__device__ __forceinline__ int wrap_axis(const DevGrid& g, int a, int v) {
    const int span = g.dmax[a] - g.dmin[a] + 1;
    if (v < g.dmin[a]) return v + span;
    if (v > g.dmax[a]) return v - span;
    return v;
}

// Boundary layer: interior nodes next to a side that is not periodic on both
// ends (their pulls resolve the boundary chain).
template <int D>
__device__ __forceinline__ bool in_shell(const DevGrid& g, const DevBC& bc, const int c[3]) {
    bool s = false;
#pragma unroll
    for (int a = 0; a < D; ++a) s |= !bc.both_periodic[a] && (c[a] == g.dmin[a] || c[a] == g.dmax[a]);
    return s;
}

template <int D>
__device__ __forceinline__ bool interior_node(const DevGrid& g, long long t, int c[3]) {
    const int nx = g.dmax[0] - g.dmin[0] + 1, ny = g.dmax[1] - g.dmin[1] + 1;
    const int nz = (D == 3) ? g.dmax[2] - g.dmin[2] + 1 : 1;
    if (t >= (long long)nx * ny * nz) return false;
    c[0] = g.dmin[0] + (int)(t % nx);
    c[1] = g.dmin[1] + (int)((t / nx) % ny);
    c[2] = g.dmin[2] + (int)(t / ((long long)nx * ny));
    return true;
}

// Deterministic block sum of 4 doubles into part[4 * blockIdx.x + k].
__device__ __forceinline__ void block_sum4(double v0, double v1, double v2, double v3, double* part) {
    __shared__ double sh[4][256];
    const int t = threadIdx.x;
    sh[0][t] = v0;
    sh[1][t] = v1;
    sh[2][t] = v2;
    sh[3][t] = v3;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (t < s)
#pragma unroll
            for (int k = 0; k < 4; ++k) sh[k][t] += sh[k][t + s];
        __syncthreads();
    }
    if (t == 0)
#pragma unroll
        for (int k = 0; k < 4; ++k) part[4 * blockIdx.x + k] = sh[k][0];
}

constexpr int ADJ_TB = 256;

// Pull source of population q at node c: c - e_q, wrapped along periodic
// axes (every source of a bulk node is interior).
template <class L>
__device__ __forceinline__ long long pull_source(const DevGrid& g, const int c[3], int q) {
    const int sx = wrap_axis(g, 0, c[0] - L::ex(q)), sy = wrap_axis(g, 1, c[1] - L::ey(q));
    const int sz = (L::D == 3) ? wrap_axis(g, 2, c[2] - L::ez(q)) : c[2];
    return g.idx(sx, sy, sz);
}

// Boundary-layer node c, direction q: zero the adjoint entries of S^n whose
// only reader it is, or which have none.  The reader of entry (y, q) is
// y + e_q wrapped along axes periodic on both ends: for y = c - e_q that is c
// itself; entry (c, q) has no reader when c + e_q leaves the domain.
template <class L, class CT>
__device__ __forceinline__ void zero_owned_entries(const AdjArgs<CT>& a, const int c[3], long long n, int q) {
    const DevGrid& g = a.g;
    int y[3] = {c[0] - L::ex(q), c[1] - L::ey(q), c[2] - L::ez(q)};
    int x[3] = {c[0] + L::ex(q), c[1] + L::ey(q), c[2] + L::ez(q)};
#pragma unroll
    for (int ax = 0; ax < L::D; ++ax)
        if (a.bc.both_periodic[ax]) {
            y[ax] = wrap_axis(g, ax, y[ax]);
            x[ax] = wrap_axis(g, ax, x[ax]);
        }
    if (g.interior(y, L::D)) {
        const long long t = (long long)q * g.nodes + g.idx(y[0], y[1], y[2]);
        a.lFo[t] = CT(0);
        a.lGo[t] = CT(0);
    }
    if (!g.interior(x, L::D)) {
        a.lFo[(long long)q * g.nodes + n] = CT(0);
        a.lGo[(long long)q * g.nodes + n] = CT(0);
    }
}

// Transpose of the bulk pull for g at bulk node c: (source of q, q) <- the
// g_post adjoint plus the phase sum's share (phi' = sum_q g_post).
template <class L, class CT>
__device__ __forceinline__ void scatter_g(const AdjArgs<CT>& a, const int c[3], long long n, CT lphi) {
#pragma unroll
    for (int q = 0; q < L::Q; ++q)
        a.lGo[(long long)q * a.g.nodes + pull_source<L>(a.g, c, q)] = a.lgp[(long long)q * a.g.nodes + n] + lphi;
}

template <class L, class CT>
__global__ __launch_bounds__(ADJ_TB) void k_adj_node(AdjArgs<CT> a) {
    constexpr int Q = L::Q, D = L::D;
    const DevGrid& g = a.g;
    int c[3];
    double pb = 0, pk = 0, pm = 0;
    if (interior_node<D>(g, (long long)blockIdx.x * blockDim.x + threadIdx.x, c)) {
        const long long n = g.idx(c[0], c[1], c[2]);
        const bool shell = in_shell<D>(g, a.bc, c);
        CT fp[Q], gp[Q], lfs[Q], lgs[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            if (shell) {  // resolved by k_adj_shell_post into lfp / lgp
                fp[q] = a.lfp[(long long)q * g.nodes + n];
                gp[q] = a.lgp[(long long)q * g.nodes + n];
            } else {
                const int sx = wrap_axis(g, 0, c[0] - L::ex(q)), sy = wrap_axis(g, 1, c[1] - L::ey(q));
                const int sz = (D == 3) ? wrap_axis(g, 2, c[2] - L::ez(q)) : c[2];
                const long long src = g.idx(sx, sy, sz);
                fp[q] = a.F[(long long)q * g.nodes + src];
                gp[q] = a.G[(long long)q * g.nodes + src];
            }
            lfs[q] = a.lF[(long long)q * g.nodes + n];
            lgs[q] = a.lG[(long long)q * g.nodes + n];
        }
        CT c0, dc[3], lap;
        phi_stencil<L, CT>(a.phin, g, n, c0, dc, lap);
        CT uo[3], lu[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            uo[d] = (d < D) ? a.uo[(long long)d * g.nodes + n] : CT(0);
            lu[d] = (d < D) ? a.lU[(long long)d * g.nodes + n] : CT(0);
        }
        NodeAdj<L, CT> o{};
        node_reverse<L, CT>(a.ph, fp, gp, c0, dc, lap, uo, lfs, lgs, lu, o, a.nfloor);
        // f_post / g_post of an outflow boundary node are also the next
        // step's outflow planes (store_prev)
        if (shell && a.bc.is_fc && a.lPf_in) {
#pragma unroll
            for (int side = 0; side < 2 * D; ++side) {
                const int ax = side >> 1;
                const int inward = (side & 1) ? -1 : 1;
                if (a.bc.type[side] != BC_FREE_CONVECT || c[ax] != ((side & 1) ? g.dmax[ax] : g.dmin[ax])) continue;
                const long long t = plane_index(g, side, c), ps = plane_size(g, side);
#pragma unroll
                for (int q = 0; q < Q; ++q)
                    if (L::e(q, ax) == inward) {
                        o.lfp[q] += a.lPf_in[a.poff[side] + (long long)q * ps + t];
                        o.lgp[q] += a.lPg_in[a.poff[side] + (long long)q * ps + t];
                    }
            }
        }
        if (!shell) {
            // transpose of the bulk pull: (source of q, q) <- this node's f_post adjoint
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                a.lFo[(long long)q * g.nodes + pull_source<L>(g, c, q)] = o.lfp[q];
                a.lgp[(long long)q * g.nodes + n] = o.lgp[q];
            }
        } else {
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                a.lfp[(long long)q * g.nodes + n] = o.lfp[q];
                a.lgp[(long long)q * g.nodes + n] = o.lgp[q];
                // the entry this node reads (q from its source), and the node's
                // own entry q when nothing reads it (q leaves the domain)
                zero_owned_entries<L>(a, c, n, q);
            }
        }
        a.lst[n] = o.lc0;
        a.lsv[n] = StAdj<CT>{{o.ldc[0], o.ldc[1], D == 3 ? o.ldc[2] : CT(0)}, o.llap};
#pragma unroll
        for (int d = 0; d < D; ++d) a.lU[(long long)d * g.nodes + n] = o.luo[d];
        pb = o.lbeta;
        pk = o.lkappa;
        pm = o.lmob;
    }
    block_sum4(pb, pk, pm, 0.0, a.part);
}

// Forward f_post / g_post of the boundary-layer nodes (the Puller chain on
// S^n) into lfp / lgp, which k_adj_node reads at those nodes and then
// overwrites with the adjoints (same thread, same entries).  blockIdx.y is q,
// so each wave resolves one compile-time direction (as k_shell_pull).
template <class L, class CT, int Q0>
__device__ __forceinline__ void adj_post_dir(const AdjArgs<CT>& a, const CT* const* pfs, const CT* const* pgs,
                                             const int c[3], long long n) {
    Puller<L, CT, CT, StorePlain<CT, CT>, false> Pf{a.F, a.g, a.bc, a.uo, pfs, nullptr};
    Puller<L, CT, CT, StorePlain<CT, CT>, false> Pg{a.G, a.g, a.bc, a.uo, pgs, nullptr};
    a.lfp[(long long)Q0 * a.g.nodes + n] = Pf.lvl3(c, Q0);
    a.lgp[(long long)Q0 * a.g.nodes + n] = Pg.lvl3(c, Q0);
}
template <class L, class CT>
__global__ __launch_bounds__(ADJ_TB) void k_adj_shell_post(AdjArgs<CT> a, ShellDesc sd) {
    const unsigned S = (unsigned)shell_total(sd);
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    int c[3];
    if (t >= S || !shell_node_at<L>(a.g, sd, t, c)) return;
    const long long n = a.g.idx(c[0], c[1], c[2]);
    const CT* pfs[6];
    const CT* pgs[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        pfs[s] = a.pf ? a.pf + a.poff[s] : nullptr;
        pgs[s] = a.pg ? a.pg + a.poff[s] : nullptr;
    }
#define CLIP_ADJ_DIR(k)                                                  \
    case k:                                                              \
        if constexpr (k < L::Q) adj_post_dir<L, CT, k>(a, pfs, pgs, c, n); \
        break;
    switch (blockIdx.y) {
        CLIP_ADJ_DIR(0) CLIP_ADJ_DIR(1) CLIP_ADJ_DIR(2) CLIP_ADJ_DIR(3) CLIP_ADJ_DIR(4) CLIP_ADJ_DIR(5)
        CLIP_ADJ_DIR(6) CLIP_ADJ_DIR(7) CLIP_ADJ_DIR(8) CLIP_ADJ_DIR(9) CLIP_ADJ_DIR(10) CLIP_ADJ_DIR(11)
        CLIP_ADJ_DIR(12) CLIP_ADJ_DIR(13) CLIP_ADJ_DIR(14) CLIP_ADJ_DIR(15) CLIP_ADJ_DIR(16) CLIP_ADJ_DIR(17)
        CLIP_ADJ_DIR(18)
        default: break;
    }
#undef CLIP_ADJ_DIR
}

// acc[k] += sum over blocks of part[4 b + k], run by a whole block
// (blockDim.x a power of two, 8..256; every thread must call it).  Thread t
// sums k = t % 4 over b = t / 4, t / 4 + blockDim.x / 4, ... and the per-thread
// sums meet in a fixed tree, so the result does not depend on timing.
__device__ __forceinline__ void accumulate4(const double* part, int blocks, double* acc) {
    __shared__ double sh[256];
    const int t = threadIdx.x;
    double s = 0.0;
    for (int b = t >> 2; b < blocks; b += (int)(blockDim.x >> 2)) s += part[4 * b + (t & 3)];
    sh[t] = s;
    __syncthreads();
    for (int w = (int)blockDim.x >> 1; w >= 4; w >>= 1) {
        if (t < w) sh[t] += sh[t + w];  // w is a multiple of 4: t and t + w share k
        __syncthreads();
    }
    if (t < 4) acc[t] += sh[t];
}

// Transpose of the stencil at every cell y of phi' (interior and ghost ring):
// lall[y] = lPhi[y] (adjoint from later steps) + the node's own phi term
// (interior y) + the stencil weights of every interior node z within one
// cell of y.  With SCATTER (a reverse step; not at the initial state) bulk
// cells then scatter their g adjoint (scatter_g), and the launch's one extra
// block, its first, folds k_adj_node's `fold` partials into acc while the
// others run.
template <class L, class CT, bool SCATTER>
__global__ __launch_bounds__(ADJ_TB) void k_adj_phi(AdjArgs<CT> a, CT* lall, double* acc, int fold) {
    constexpr int D = L::D;
    const DevGrid& g = a.g;
    if (SCATTER && blockIdx.x == 0) {
        accumulate4(a.part, fold, acc);
        return;
    }
    const long long y = (long long)(blockIdx.x - (SCATTER ? 1 : 0)) * blockDim.x + threadIdx.x;
    int c[3];
    if (y >= g.nodes || !g.coords(y, c)) return;
    CT total = a.lPhi[y];
    if (g.interior(c, D)) total += a.lst[y];
    const int zlo = (D == 3) ? -1 : 0, zhi = (D == 3) ? 1 : 0;
    for (int dz = zlo; dz <= zhi; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int z[3] = {c[0] - dx, c[1] - dy, c[2] - dz};
                if (!g.interior(z, D)) continue;
                CT wg[3], wl;
                stencil_weights<D, CT>(dx, dy, dz, wg, wl);
                const StAdj<CT> r = a.lsv[g.idx(z[0], z[1], z[2])];
                CT v = wl * r.lap;
#pragma unroll
                for (int d = 0; d < D; ++d) v += wg[d] * r.dc[d];
                total += v;
            }
    lall[y] = total;
    if (SCATTER && g.interior(c, D) && !in_shell<D>(g, a.bc, c)) scatter_g<L>(a, c, y, total);
}

// Fully periodic domains: the ring of phi' is the periodic wrap of phi^n's
// interior, so the stencil transpose gathers with wrapped reads in one
// kernel at interior nodes: lphin = the adjoint of phi' and, with LAG, the
// reads that hit the lagged ring go to lPhi (the adjoint of phi^n); without
// LAG (the initial state, whose ring wraps phi0) they stay in lphin.  A
// reverse step (LAG) then scatters each node's g adjoint (scatter_g), and the
// launch's one extra block, its first, folds k_adj_node's `fold` partials
// into acc while the others run.
template <class L, class CT, bool LAG>
__global__ __launch_bounds__(ADJ_TB) void k_adj_phi_periodic(AdjArgs<CT> a, double* acc, int fold) {
    constexpr int D = L::D;
    const DevGrid& g = a.g;
    if (LAG && blockIdx.x == 0) {
        accumulate4(a.part, fold, acc);
        return;
    }
    int c[3];
    if (!interior_node<D>(g, (long long)(blockIdx.x - (LAG ? 1 : 0)) * blockDim.x + threadIdx.x, c)) return;
    const long long n = g.idx(c[0], c[1], c[2]);
    CT total = a.lPhi[n] + a.lst[n];
    CT lag = CT(0);
    const int zlo = (D == 3) ? -1 : 0, zhi = (D == 3) ? 1 : 0;
    for (int dz = zlo; dz <= zhi; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                CT wg[3], wl;
                stencil_weights<D, CT>(dx, dy, dz, wg, wl);
                const int z[3] = {c[0] - dx, c[1] - dy, c[2] - dz};
                const bool zin = g.interior(z, D);
                const int zx = wrap_axis(g, 0, z[0]), zy = wrap_axis(g, 1, z[1]);
                const int zz = (D == 3) ? wrap_axis(g, 2, z[2]) : z[2];
                const StAdj<CT> r = a.lsv[g.idx(zx, zy, zz)];
                CT v = wl * r.lap;
#pragma unroll
                for (int d = 0; d < D; ++d) v += wg[d] * r.dc[d];
                if (zin || !LAG) total += v;
                else lag += v;
            }
    a.lphin[n] = total;
    if (LAG) {
        a.lPhi[n] = lag;
        scatter_g<L>(a, c, n, total);
    }
}

// Transpose of a copy map (host-built rows: target = a cell of the source
// field, entries' rd = the cells copied from it): out[target] = sum over the
// row of in[rd].  Targets of no row are left as they are.
template <class CT>
__global__ __launch_bounds__(ADJ_TB) void k_adj_ring(AdjRows r, const CT* in, CT* out) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= r.rows) return;
    CT s = CT(0);
    for (int k = r.start[row]; k < r.start[row + 1]; ++k) s += in[r.e[k].rd];
    out[r.target[row]] += s;
}

template <class CT>
__global__ void k_adj_fill(CT* x, long long n, CT v) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

// Node of element t of an FP64 interior array: x fastest (t = i + nx (j +
// ny k)) or x slowest (t = (i ny + j) nz + k, a C array [nx][ny][nz]); the
// interior extents are dmax - dmin + 1 per axis (1 on z in 2D).
__device__ __forceinline__ long long interior_node(const DevGrid& g, long long t, int xslow) {
    const long long n0 = g.dmax[0] - g.dmin[0] + 1, n1 = g.dmax[1] - g.dmin[1] + 1, n2 = g.dmax[2] - g.dmin[2] + 1;
    long long i, j, k;
    if (xslow) {
        k = t % n2;
        j = (t / n2) % n1;
        i = t / (n2 * n1);
    } else {
        i = t % n0;
        j = (t / n0) % n1;
        k = t / (n0 * n1);
    }
    return g.idx((int)i + g.dmin[0], (int)j + g.dmin[1], (int)k + g.dmin[2]);
}

// Interior array -> node array (the other nodes are left as they are).
template <class CT>
__global__ void k_adj_from_interior(CT* dst, const double* src, DevGrid g, long long count, int xslow) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) dst[interior_node(g, t, xslow)] = (CT)src[t];
}

// Node arrays -> interior array: a (+ b where b is given) at each interior node.
template <class CT>
__global__ void k_adj_to_interior(double* dst, const CT* a, const CT* b, DevGrid g, long long count, int xslow) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const long long n = interior_node(g, t, xslow);
    dst[t] = b ? (double)a[n] + (double)b[n] : (double)a[n];
}


// Weight of a boundary-chain term (1, or a free-convect weight from u^n).
template <class CT>
__device__ __forceinline__ CT chain_weight(const AdjEntry& e, const CT* u) {
    if (e.uidx < 0) return CT(1);
    const CT un = fabs(u[e.uidx]);
    return (e.branch == 0 ? un : CT(1)) / (CT(1) + un);
}

// Adjoint reaching one row's target from its boundary-layer readers: f and g
// parts (g with the phase sum's phi adjoint at the reader).
template <class CT>
__device__ __forceinline__ void row_sums(const AdjArgs<CT>& a, const AdjRows& r, int row, CT& sf, CT& sg) {
    sf = CT(0);
    sg = CT(0);
    for (int k = r.start[row]; k < r.start[row + 1]; ++k) {
        const AdjEntry& e = r.e[k];
        const CT w = chain_weight(e, a.uo);
        sf += w * a.lfp[e.rd];
        sg += w * (a.lgp[e.rd] + a.lphin[e.rd % a.g.nodes]);
    }
}

// Boundary-layer readers of state populations: lF / lG += their adjoint.
template <class CT>
__global__ __launch_bounds__(ADJ_TB) void k_adj_csr_pop(AdjArgs<CT> a, AdjRows r) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= r.rows) return;
    CT sf, sg;
    row_sums(a, r, row, sf, sg);
    a.lFo[r.target[row]] += sf;
    a.lGo[r.target[row]] += sg;
}

// Readers of outflow-plane entries: the adjoint of P^n (every entry has a
// row; the target indexes the concatenated plane sets).
template <class CT>
__global__ __launch_bounds__(ADJ_TB) void k_adj_csr_prev(AdjArgs<CT> a, AdjRows r) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= r.rows) return;
    CT sf, sg;
    row_sums(a, r, row, sf, sg);
    a.lPf_out[r.target[row]] = sf;
    a.lPg_out[r.target[row]] = sg;
}

// Readers of inlet ghost values: accumulated over the sweep (row-indexed).
template <class CT>
__global__ __launch_bounds__(ADJ_TB) void k_adj_csr_inlet(AdjArgs<CT> a, AdjRows r, double* accf, double* accg) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= r.rows) return;
    CT sf, sg;
    row_sums(a, r, row, sf, sg);
    accf[row] += (double)sf;
    accg[row] += (double)sg;
}

// u^n through the free-convect weights: d/du [(P + |u| X) / (1 + |u|)] =
// sign(u) (X - P) / (1 + |u|)^2, for f and g readers.
template <class CT>
__global__ __launch_bounds__(ADJ_TB) void k_adj_csr_u(AdjArgs<CT> a, AdjRows r) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= r.rows) return;
    const long long ui = r.target[row];
    const CT uu = a.uo[ui];
    const CT un = fabs(uu);
    const CT sg = uu > CT(0) ? CT(1) : (uu < CT(0) ? CT(-1) : CT(0));
    const CT k = sg / ((CT(1) + un) * (CT(1) + un));
    CT acc = CT(0);
    for (int j = r.start[row]; j < r.start[row + 1]; ++j) {
        const AdjEntry& e = r.e[j];
        const CT xf = a.F[e.xidx], xg = a.G[e.xidx];
        const CT pf = a.pf[a.poff[e.side] + e.pidx], pg = a.pg[a.poff[e.side] + e.pidx];
        acc += k * ((xf - pf) * a.lfp[e.rd] + (xg - pg) * (a.lgp[e.rd] + a.lphin[e.rd % a.g.nodes]));
    }
    a.lU[ui] += acc;
}

// Reverse of the initial state and first collision at every interior node
// (phin = phi0 with its periodic ghost ring).
template <class L, class CT>
__global__ __launch_bounds__(ADJ_TB) void k_adj_init(AdjArgs<CT> a) {
    constexpr int Q = L::Q, D = L::D;
    const DevGrid& g = a.g;
    int c[3];
    double pb = 0, pk = 0, pm = 0, ps = 0;
    if (interior_node<D>(g, (long long)blockIdx.x * blockDim.x + threadIdx.x, c)) {
        const long long n = g.idx(c[0], c[1], c[2]);
        CT lfs[Q], lgs[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            lfs[q] = a.lF[(long long)q * g.nodes + n];
            lgs[q] = a.lG[(long long)q * g.nodes + n];
        }
        CT c0, dc[3], lap;
        phi_stencil<L, CT>(a.phin, g, n, c0, dc, lap);
        // adjoint of the first outflow planes (the initial f / g on outflow
        // boundary nodes, k_init's store_prev)
        CT lpf[Q], lpg[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) lpf[q] = lpg[q] = CT(0);
        if (a.bc.is_fc && a.lPf_in) {
#pragma unroll
            for (int side = 0; side < 2 * D; ++side) {
                const int ax = side >> 1;
                const int inward = (side & 1) ? -1 : 1;
                if (a.bc.type[side] != BC_FREE_CONVECT || c[ax] != ((side & 1) ? g.dmax[ax] : g.dmin[ax])) continue;
                const long long t = plane_index(g, side, c), ps = plane_size(g, side);
#pragma unroll
                for (int q = 0; q < Q; ++q)
                    if (L::e(q, ax) == inward) {
                        lpf[q] += a.lPf_in[a.poff[side] + (long long)q * ps + t];
                        lpg[q] += a.lPg_in[a.poff[side] + (long long)q * ps + t];
                    }
            }
        }
        CT lc0 = CT(0), ldc[3] = {CT(0), CT(0), CT(0)}, llap = CT(0), lb = CT(0), lk = CT(0), lm = CT(0),
           lsg = CT(0);
        init_reverse<L, CT>(a.ph, a.sigma, a.half_length, c0, dc, lap, lfs, lgs, lc0, ldc, llap, lb, lk, lm, lsg,
                            a.nfloor, lpf, lpg);
        a.lst[n] = lc0;
        a.lsv[n] = StAdj<CT>{{ldc[0], ldc[1], D == 3 ? ldc[2] : CT(0)}, llap};
        pb = lb;
        pk = lk;
        pm = lm;
        ps = lsg;
    }
    block_sum4(pb, pk, pm, ps, a.part);
}

// acc[k] += sum over blocks of part[4 b + k] (one block of ADJ_TB threads).
template <class CT>
__global__ __launch_bounds__(ADJ_TB) void k_adj_accumulate(const double* part, int blocks, double* acc) {
    accumulate4(part, blocks, acc);
}

}  // namespace clipamd
