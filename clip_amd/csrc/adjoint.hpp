// adjoint.hpp — reverse-mode (adjoint) kernels of the coupled step on fully
// periodic domains, FP64 or FP32 populations stored plainly.
//
// Forward step n (state S^n = f*, g*, u, phi with its ghost ring):
//   f_post/g_post(x) = f*/g*(x - e_q)                (pull with periodic wrap)
//   phi'(x)          = sum_q g_post(x)               (k_phase)
//   stencil(x)       reads phi' at interior neighbours and phi^n (the lagged
//                    ghost ring) where the neighbour is a ghost cell
//   node(x)          u, g*, f* from f_post, g_post, phi', grad, lap, u^n
// The reverse sweep of one step runs three kernels:
//   k_adj_node  node VJP (node.hpp) at every interior node
//   k_adj_phi   transpose of the stencil: adjoint of phi' and of the lagged phi^n
//   k_adj_pull  transpose of the pull: adjoint of f*, g* of S^n
#pragma once

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

template <class CT>
struct AdjArgs {
    DevGrid g;
    Phys<CT> ph;
    const CT* F;     // f* of S^n (StorePlain), Q x nodes
    const CT* G;     // g* of S^n
    const CT* phin;  // phi^{n+1} with the ghost ring of phi^n (or phi0 at init)
    const CT* uo;    // u^n
    CT* lF;          // in: adjoint of f*^{n+1}; out: adjoint of f*^n
    CT* lG;          // in: adjoint of g*^{n+1}; out: adjoint of g*^n
    CT* lU;          // in: adjoint of u^{n+1};  out: adjoint of u^n
    CT* lPhi;        // in: adjoint of phi^{n+1}; out: adjoint of phi^n (interior)
    CT* lfp;         // scratch Q x nodes: adjoint of f_post
    CT* lgp;         // scratch Q x nodes: adjoint of g_post
    CT* lst;         // scratch (D+2) x nodes: adjoint of phi, grad phi, lap phi at the node
    CT* lphin;       // scratch nodes: adjoint of phi^{n+1}
    double* part;    // 4 x blocks partial sums: beta, kappa, mobility, sigma
    CT sigma, half_length;
    CT nfloor;       // regularisation of the normal's derivative (0 = exact)
};

__device__ __forceinline__ int wrap_axis(const DevGrid& g, int a, int v) {
    const int span = g.dmax[a] - g.dmin[a] + 1;
    if (v < g.dmin[a]) return v + span;
    if (v > g.dmax[a]) return v - span;
    return v;
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

template <class L, class CT>
__global__ __launch_bounds__(ADJ_TB) void k_adj_node(AdjArgs<CT> a) {
    constexpr int Q = L::Q, D = L::D;
    const DevGrid& g = a.g;
    int c[3];
    double pb = 0, pk = 0, pm = 0;
    if (interior_node<D>(g, (long long)blockIdx.x * blockDim.x + threadIdx.x, c)) {
        const long long n = g.idx(c[0], c[1], c[2]);
        CT fp[Q], gp[Q], lfs[Q], lgs[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int sx = wrap_axis(g, 0, c[0] - L::ex(q)), sy = wrap_axis(g, 1, c[1] - L::ey(q));
            const int sz = (D == 3) ? wrap_axis(g, 2, c[2] - L::ez(q)) : c[2];
            const long long src = g.idx(sx, sy, sz);
            fp[q] = a.F[(long long)q * g.nodes + src];
            gp[q] = a.G[(long long)q * g.nodes + src];
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
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            a.lfp[(long long)q * g.nodes + n] = o.lfp[q];
            a.lgp[(long long)q * g.nodes + n] = o.lgp[q];
        }
        a.lst[n] = o.lc0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            a.lst[(long long)(1 + d) * g.nodes + n] = o.ldc[d];
            a.lU[(long long)d * g.nodes + n] = o.luo[d];
        }
        a.lst[(long long)(1 + D) * g.nodes + n] = o.llap;
        pb = o.lbeta;
        pk = o.lkappa;
        pm = o.lmob;
    }
    block_sum4(pb, pk, pm, 0.0, a.part);
}

// Transpose of the stencil at interior node y.  LAG: the stencil of the step
// read phi^n at ghost positions (adjoint goes to lPhi); otherwise (initial
// state) every read is phi0 at the wrapped position.
template <class L, class CT, bool LAG>
__global__ __launch_bounds__(ADJ_TB) void k_adj_phi(AdjArgs<CT> a) {
    constexpr int D = L::D;
    const DevGrid& g = a.g;
    int c[3];
    if (!interior_node<D>(g, (long long)blockIdx.x * blockDim.x + threadIdx.x, c)) return;
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
                const long long m = g.idx(zx, zy, zz);
                CT v = wl * a.lst[(long long)(1 + D) * g.nodes + m];
#pragma unroll
                for (int d = 0; d < D; ++d) v += wg[d] * a.lst[(long long)(1 + d) * g.nodes + m];
                if (zin || !LAG) total += v;
                else lag += v;
            }
    a.lphin[n] = total;
    if (LAG) a.lPhi[n] = lag;
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

// Transpose of the pull: adjoint of f*, g* of S^n at interior node y.

template <class L, class CT>
__global__ __launch_bounds__(ADJ_TB) void k_adj_pull(AdjArgs<CT> a, double* acc, int blocks) {
    constexpr int Q = L::Q, D = L::D;
    const DevGrid& g = a.g;
    // the launch has one block more than the pull needs: that block folds
    // k_adj_node's partials, next to the pull rather than after it
    if (blockIdx.x == (unsigned)blocks) {
        accumulate4(a.part, blocks, acc);
        return;
    }
    int c[3];
    if (!interior_node<D>(g, (long long)blockIdx.x * blockDim.x + threadIdx.x, c)) return;
    const long long n = g.idx(c[0], c[1], c[2]);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int xx = wrap_axis(g, 0, c[0] + L::ex(q)), xy = wrap_axis(g, 1, c[1] + L::ey(q));
        const int xz = (D == 3) ? wrap_axis(g, 2, c[2] + L::ez(q)) : c[2];
        const long long x = g.idx(xx, xy, xz);
        a.lF[(long long)q * g.nodes + n] = a.lfp[(long long)q * g.nodes + x];
        a.lG[(long long)q * g.nodes + n] = a.lgp[(long long)q * g.nodes + x] + a.lphin[x];
    }
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
        CT lc0 = CT(0), ldc[3] = {CT(0), CT(0), CT(0)}, llap = CT(0), lb = CT(0), lk = CT(0), lm = CT(0),
           lsg = CT(0);
        init_reverse<L, CT>(a.ph, a.sigma, a.half_length, c0, dc, lap, lfs, lgs, lc0, ldc, llap, lb, lk, lm, lsg,
                            a.nfloor);
        a.lst[n] = lc0;
#pragma unroll
        for (int d = 0; d < D; ++d) a.lst[(long long)(1 + d) * g.nodes + n] = ldc[d];
        a.lst[(long long)(1 + D) * g.nodes + n] = llap;
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
