// kernels.hpp — HIP device code of the coupled Allen–Cahn / WMRT step (gfx950).
//
// Layout in HBM: structure of arrays, x fastest.  Node (i, j, k) of the
// ghost-inclusive grid has index n = i + j*ext_x + k*ext_x*ext_y, population q
// of buffer A sits at A[q*nodes + n].  f and g live in two ping-pong buffers
// each (post-collision values, ghost cells included); phi in two (the phi of
// the collision that wrote the buffer being pulled, and the new one, each with
// its ghost ring); u in two (old / new).
//
// One time step is three kernel classes:
//   k_ghost_*   refresh the new phi's ghost ring from the old phi (periodic
//               wrap, inlet, mirror): the lagged ghost values the reference's
//               stencils see
//   k_phase     phi^{n+1} = sum_q g_post, g_post pulled from g* with every
//               boundary fix-up resolved at pull time
//   k_fused     pull f_post/g_post, phi stencil (mu, grad phi, normal),
//               macroscopic p/u (from u_old), then both collisions; writes
//               f*, g* into the other buffer and u into the other u buffer.
// k_phase and k_fused run as a bulk launch (plain offset pulls with periodic
// wrap) and a boundary-layer launch (nodes next to a non-periodic side).
// Reference phases folded into these kernels (NSAllen::solve, NsAllen.cu:881-892):
// normal_FD, kernelCollisionMRTg/f, kernelPeriodicBoundary, JetBoundary,
// kernelStreaming, kernelHalfBounceBack, kernelFreeConvect, kernelNeumann,
// kernelMirrorBoundary, kernelMacroscopicg, Chemical_Potential,
// Isotropic_Gradient, kernelMacroscopicf.
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <type_traits>

#include "node.hpp"

namespace clipamd {

enum : int {
    BC_WALL = CLIP_BC_WALL,
    BC_SLIP_WALL = CLIP_BC_SLIP_WALL,
    BC_FREE_CONVECT = CLIP_BC_FREE_CONVECT,
    BC_PERIODIC = CLIP_BC_PERIODIC,
    BC_NEUMANN = CLIP_BC_NEUMANN,
    BC_VELOCITY = CLIP_BC_VELOCITY,
    BC_DO_NOTHING = CLIP_BC_DO_NOTHING
};

// Device layout of a node field: node (i, j, k) at i + j*sy + k*sz.  The row
// stride sy may exceed ext[0] (padding cells x >= ext[0] belong to no node),
// so that with the buffer offsets of the solver every row's first interior
// node starts a 128-byte line; nodes = sz * ext[2] is the per-field stride.
struct DevGrid {
    int ext[3];
    int dmin[3], dmax[3];
    long long sy, sz;  // strides of j and k (i has stride 1)
    long long nodes;
    __host__ __device__ long long idx(int i, int j, int k) const { return i + j * sy + k * sz; }
    // Coordinates of device index n; false for a padding cell.
    __host__ __device__ bool coords(long long n, int c[3]) const {
        c[2] = (int)(n / sz);
        const long long r = n - (long long)c[2] * sz;
        c[1] = (int)(r / sy);
        c[0] = (int)(r - (long long)c[1] * sy);
        return c[0] < ext[0];
    }
    __host__ __device__ bool interior(const int c[3], int D) const {
        bool in = c[0] >= dmin[0] && c[0] <= dmax[0] && c[1] >= dmin[1] && c[1] <= dmax[1];
        if (D == 3) in = in && c[2] >= dmin[2] && c[2] <= dmax[2];
        return in;
    }
};

struct DevBC {
    int type[6];
    int periodic[3];       // axis wraps (either side periodic: Solver.cu:96, 110, 125)
    int both_periodic[3];  // both sides periodic: no boundary-layer fix-ups on this axis
    int is_wall, is_fc, is_neumann, is_velocity, mirror_gate;
    int slip_specular;  // slip sides reflect specularly (opt-in; see wall_source)
    int wall_pass;      // the half-way wall pass runs: any wall side, or a specular slip side
    const unsigned char* vflag[6];  // per velocity side: 1 = ghost node inside the inlet disc
};

// Grid of a configuration: CLIP's extents (2D: N + 2 with the ghost ring
// outside N; 3D: N with the ring inside N, Domain.cu:22-82), interior
// [1, ext - 2] per axis, and the row stride rounded up to a multiple of
// `unit` elements (1: dense).
inline DevGrid grid_of_config(const clip_config& c, int D, long long unit) {
    DevGrid g{};
    for (int a = 0; a < 3; ++a) {
        if (D == 2) {
            g.ext[a] = (a < 2) ? (int)c.N[a] + 2 : 1;
            g.dmin[a] = (a < 2) ? 1 : 0;
            g.dmax[a] = (a < 2) ? g.ext[a] - 2 : 0;
        } else {
            g.ext[a] = (int)c.N[a];
            g.dmin[a] = 1;
            g.dmax[a] = g.ext[a] - 2;
        }
    }
    g.sy = (g.ext[0] + unit - 1) / unit * unit;
    g.sz = g.sy * g.ext[1];
    g.nodes = g.sz * g.ext[2];
    return g;
}

// Side types and the launch flags of the reference's boundary classes
// (Boundary.cu:291-318); the inlet disc flags are set by the solver.
inline DevBC bc_of_config(const clip_config& c, int D) {
    DevBC bc{};
    for (int s = 0; s < 6; ++s) bc.type[s] = (s < 2 * D) ? c.bc_type[s] : -1;
    bool isWall = false, isFC = false, isNeu = false, isVel = false, isSlip = false;
    for (int s = 0; s < 2 * D; ++s) {
        isWall |= c.bc_type[s] == CLIP_BC_WALL;
        isSlip |= c.bc_type[s] == CLIP_BC_SLIP_WALL;
        isFC |= c.bc_type[s] == CLIP_BC_FREE_CONVECT;
        isNeu |= c.bc_type[s] == CLIP_BC_NEUMANN;
        isVel |= c.bc_type[s] == CLIP_BC_VELOCITY;
    }
    for (int a = 0; a < 3; ++a) {
        const bool lo = a < D && c.bc_type[2 * a] == CLIP_BC_PERIODIC;
        const bool hi = a < D && c.bc_type[2 * a + 1] == CLIP_BC_PERIODIC;
        bc.periodic[a] = lo || hi;
        bc.both_periodic[a] = lo && hi;
    }
    bc.is_wall = isWall;
    bc.is_fc = isFC;
    bc.is_neumann = isNeu;
    bc.is_velocity = isVel;
    bc.mirror_gate = isWall || isFC || isSlip || isNeu;
    bc.slip_specular = c.slip_specular != 0;
    bc.wall_pass = isWall || (bc.slip_specular && isSlip);
    return bc;
}

// Half-way wall pass (kernelHalfBounceBack, Solver.cu:248-333) at interior
// node c for incoming direction q: the direction of the same node's
// post-collision population that becomes post-streaming population q, or -1
// when no wall-type side of c takes q in.
//  - Reference rule (slip_specular = 0; the pass runs when any side is a
//    wall): wall and slip sides both bounce back, q <- opp(q) (wallBCMap).
//  - Specular rule (slip_specular = 1; the pass runs when any side is a wall
//    or a slip wall): wall sides bounce back through wallBCMap, slip sides
//    reflect through slipWallBCMap (WMRT.cuh:243-264), q <- q mirrored along
//    the side's axis.  On a node shared by several such sides the assignment
//    the reference kernel's loop makes last (list position m outer, sides
//    X-, X+, Y-, Y+, Z-, Z+ inner) stands.
template <class L>
__host__ __device__ inline int wall_source(const DevGrid& g, const DevBC& bc, const int c[3], int q) {
    if (!bc.wall_pass) return -1;
    int best = -1, src = -1;
    for (int side = 0; side < 2 * L::D; ++side) {
        const int a = side >> 1;
        const int inward = (side & 1) ? -1 : 1;
        const int t = bc.type[side];
        if (!(t == BC_WALL || t == BC_SLIP_WALL) || L::e(q, a) != inward ||
            c[a] != ((side & 1) ? g.dmax[a] : g.dmin[a]))
            continue;
        if (!bc.slip_specular) return L::opp(q);
        const bool slip = t == BC_SLIP_WALL;
        const int key = (slip ? L::slip_pos(side, q) : L::wall_pos(side, q)) * 6 + side;
        if (key > best) {
            best = key;
            src = slip ? L::mirror(q, a) : L::opp(q);
        }
    }
    return src;
}

// Index of a ghost-plane (or boundary-plane) cell of side `side` within its
// plane: the two tangential coordinates, lower axis fastest.
__host__ __device__ inline long long plane_index(const DevGrid& g, int side, const int c[3]) {
    const int a = side >> 1;
    const int t1 = (a == 0) ? 1 : 0;
    const int t2 = (a == 2) ? 1 : 2;
    return (long long)c[t2] * g.ext[t1] + c[t1];
}
__host__ __device__ inline long long plane_size(const DevGrid& g, int side) {
    const int a = side >> 1;
    const int t1 = (a == 0) ? 1 : 0;
    const int t2 = (a == 2) ? 1 : 2;
    return (long long)g.ext[t1] * g.ext[t2];
}


// ------------------------------------------------------------ storage
// Storage policies convert between the stored population and the arithmetic
// type.  `ref` is the phase field of the node whose collision wrote the value
// (0 for ghost cells and for f); the plain policies ignore it.
template <class ST, class CT>
struct StorePlain {
    static constexpr bool uses_ref = false;
    __device__ static CT ld(const ST* p, long long i, int /*q*/, CT /*ref*/) { return (CT)p[i]; }
    __device__ static void st(ST* p, long long i, int /*q*/, CT v, CT /*ref*/) { p[i] = (ST)v; }
};

// 16-bit storage of FP32 populations as scaled deviations from the lattice-
// weight equilibrium at rest: g is stored as (g - w_q phi_ref) * 2^10, f as
// f * 2^10 (its rest equilibrium w_q p is ~0: p is the dynamic pressure).
template <class L, class ST>
struct StoreMixed {
    static constexpr bool uses_ref = true;
    static constexpr float kScale = 1024.0f;
    __device__ static float to_f(__half h) { return __half2float(h); }
    __device__ static float to_f(__hip_bfloat16 h) { return __bfloat162float(h); }
    __device__ static void from_f(__half& h, float v) { h = __float2half_rn(v); }
    __device__ static void from_f(__hip_bfloat16& h, float v) { h = __float2bfloat16(v); }
    __device__ static float ld(const ST* p, long long i, int q, float ref) {
        return to_f(p[i]) * (1.0f / kScale) + float(L::w(q)) * ref;
    }
    __device__ static void st(ST* p, long long i, int q, float v, float ref) {
        from_f(p[i], (v - float(L::w(q)) * ref) * kScale);
    }
};

// ------------------------------------------------------------ arguments
// Boundary layer = interior nodes on a plane next to a non-periodic side.
// The planes of the first such axis are taken whole, those of later axes
// exclude the rows already covered.  `axes` ascending, `count` = nodes of
// each axis' plane pair.
struct ShellDesc {
    int na;
    int axes[3];
    long long count[3];
};
__host__ __device__ inline long long shell_total(const ShellDesc& sd) {
    long long t = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) t += (k < sd.na) ? sd.count[k] : 0;  // constant indices
    return t;
}

template <class ST, class CT>
struct StepArgs {
    DevGrid g;
    DevBC bc;
    ShellDesc shell;
    Phys<CT> ph;
    const ST* fin;   // f* of the buffer being pulled
    const ST* gin;
    ST* fout;        // f* written by the collision
    ST* gout;
    CT* phi;            // new phi with ghost ring (k_phase writes, k_fused reads the stencil)
    const CT* phi_ref;  // phi the pulled buffer was encoded with (nullptr = 0)
    const CT* uin;   // u of the previous macroscopic step (D planes): u_old of the velocity
    const CT* ufc;   // u the outflow fix-ups of the pull read (the u of the streaming step's start)
    CT* uout;        // u of this macroscopic step
    const CT* ucol;  // u the collide-only mode collides with
    CT* p;           // pressure (read by the collide-only mode; written by exports)
    CT* shell_f;     // boundary-layer f_post / g_post (Q x shell nodes), written by k_shell_pull
    CT* shell_g;
    const CT* prev_f_in[6];  // outflow planes: f_post of the previous step
    CT* prev_f_out[6];
    const CT* prev_g_in[6];
    CT* prev_g_out[6];
};

template <class SP, class CT>
__device__ __forceinline__ CT gref(const CT* phi_ref, long long node) {
    if constexpr (SP::uses_ref) return phi_ref ? phi_ref[node] : CT(0);
    else return CT(0);
}

// Reference node of the stored population q of interior node s: the node
// that pulls it next, s + e_q, wrapped along axes periodic on both sides.
// The mixed policy stores g relative to w_q phi there, so a bulk node
// decodes all Q pulled values with its own phi (one load).
template <class L>
__device__ __forceinline__ long long ref_index(const DevGrid& g, const DevBC& bc, const int s[3], int q) {
    int d[3] = {s[0] + L::ex(q), s[1] + L::ey(q), s[2] + L::ez(q)};
#pragma unroll
    for (int a = 0; a < L::D; ++a) {
        if (!bc.both_periodic[a]) continue;
        const int span = g.dmax[a] - g.dmin[a] + 1;
        if (d[a] < g.dmin[a]) d[a] += span;
        else if (d[a] > g.dmax[a]) d[a] -= span;
    }
    return g.idx(d[0], d[1], d[2]);
}

// ------------------------------------------------------------ pull with fix-ups
// Post-streaming population q at interior node c, reproducing the reference's
// streaming + boundary sequence:
//   raw   kernelStreaming from the ghost-updated f (periodic wrap, inlet)
//   lvl1  kernelHalfBounceBack (wall / slip wall, when any wall exists)
//   lvl2  kernelFreeConvect (reads lvl1 of the inner neighbour, Jacobi)
//   lvl3  kernelNeumann     (reads lvl2 of the inner neighbour)
// GREF: the distribution is stored relative to phi_ref (g of the mixed policy).
template <class L, class ST, class CT, class SP, bool GREF>
struct Puller {
    const ST* A;
    const DevGrid& g;
    const DevBC& bc;
    const CT* uold;
    const CT* const* prev;
    const CT* phi_ref;

    __device__ __forceinline__ CT load(int q, const int s[3]) const {
        const long long node = g.idx(s[0], s[1], s[2]);
        const CT r = (GREF && g.interior(s, L::D)) ? gref<SP>(phi_ref, ref_index<L>(g, bc, s, q)) : CT(0);
        return SP::ld(A, (long long)q * g.nodes + node, q, r);
    }

    __device__ __forceinline__ CT raw(const int c[3], int q) const {
        int s[3] = {c[0] - L::ex(q), c[1] - L::ey(q), c[2] - L::ez(q)};
#pragma unroll
        for (int a = 0; a < L::D; ++a) {
            if (bc.periodic[a]) {
                if (s[a] < g.dmin[a]) s[a] = g.dmax[a];
                else if (s[a] > g.dmax[a]) s[a] = g.dmin[a];
            }
        }
        if (bc.is_velocity) {
#pragma unroll
            for (int side = 2 * L::D - 1; side >= 0; --side) {
                const int a = side >> 1;
                const int inward = (side & 1) ? -1 : 1;
                if (bc.type[side] != BC_VELOCITY || L::e(q, a) != inward) continue;
                const int ghost = (side & 1) ? g.dmax[a] + 1 : g.dmin[a] - 1;
                if (s[a] != ghost) continue;
                if (!bc.vflag[side][plane_index(g, side, s)]) {
                    // outside the inlet disc: bounce-back copy from the
                    // boundary node next to the ghost (Solver.cu:621-633)
                    int b[3] = {s[0], s[1], s[2]};
                    b[a] = (side & 1) ? g.dmax[a] : g.dmin[a];
                    return load(L::opp(q), b);
                }
                break;  // inside: the time-invariant inlet value stored in the ghost cell
            }
        }
        return load(q, s);
    }

    __device__ __forceinline__ CT lvl1(const int c[3], int q) const {
        if (bc.wall_pass) {
            const int src = wall_source<L>(g, bc, c, q);
            if (src >= 0) return load(src, c);
        }
        return raw(c, q);
    }

    __device__ __forceinline__ CT lvl2(const int c[3], int q) const {
        if (bc.is_fc) {
#pragma unroll
            for (int side = 2 * L::D - 1; side >= 0; --side) {
                const int a = side >> 1;
                const int inward = (side & 1) ? -1 : 1;
                if (bc.type[side] != BC_FREE_CONVECT || L::e(q, a) != inward ||
                    c[a] != ((side & 1) ? g.dmax[a] : g.dmin[a]))
                    continue;
                int nb[3] = {c[0], c[1], c[2]};
                nb[a] += inward;
                const CT cv = fabs(uold[(long long)a * g.nodes + g.idx(nb[0], nb[1], nb[2])]);
                const CT prv = prev[side][(long long)q * plane_size(g, side) + plane_index(g, side, c)];
                return (prv + cv * lvl1(nb, q)) / (CT(1) + cv);
            }
        }
        return lvl1(c, q);
    }

    __device__ __forceinline__ CT lvl3(const int c[3], int q) const {
        if (bc.is_neumann) {
#pragma unroll
            for (int side = 2 * L::D - 1; side >= 0; --side) {
                const int a = side >> 1;
                const int inward = (side & 1) ? -1 : 1;
                if (bc.type[side] != BC_NEUMANN || L::e(q, a) != inward ||
                    c[a] != ((side & 1) ? g.dmax[a] : g.dmin[a]))
                    continue;
                int nb[3] = {c[0], c[1], c[2]};
                nb[a] += inward;
                return lvl2(nb, q);
            }
        }
        return lvl2(c, q);
    }
};

template <class L>
__device__ inline bool on_edge(const DevGrid& g, const int c[3]) {
    bool e = c[0] == g.dmin[0] || c[0] == g.dmax[0] || c[1] == g.dmin[1] || c[1] == g.dmax[1];
    if (L::D == 3) e = e || c[2] == g.dmin[2] || c[2] == g.dmax[2];
    return e;
}

// Writes the post-streaming value of an outflow boundary node into the
// outflow plane of the next step (f_prev / g_prev, NsAllen.cu:850-854).
template <class L, class CT>
__device__ inline void store_prev(const DevGrid& g, const DevBC& bc, CT* const* prev_out, const int c[3], int q,
                                  CT v) {
#pragma unroll
    for (int side = 0; side < 2 * L::D; ++side) {
        const int a = side >> 1;
        const int inward = (side & 1) ? -1 : 1;
        if (bc.type[side] == BC_FREE_CONVECT && L::e(q, a) == inward &&
            c[a] == ((side & 1) ? g.dmax[a] : g.dmin[a]))
            prev_out[side][(long long)q * plane_size(g, side) + plane_index(g, side, c)] = v;
    }
}

// ------------------------------------------------------------ node math
// Division by a compile-time constant: exact quotient in FP64, multiplication
// by the reciprocal in FP32.
template <class CT>
__device__ __forceinline__ CT cdiv(CT x, double d) {
    if constexpr (sizeof(CT) == 8) return div_rn(x, d, 1.0 / d);
    else return x * CT(1.0 / d);
}

// Velocity part of the equilibrium, Equilibrium_new (NsAllen.cu:77-94).
template <class L, class CT>
__device__ __forceinline__ CT gamma_u(int q, const CT u[3]) {
    CT eu = CT(L::ex(q)) * u[0] + CT(L::ey(q)) * u[1];
    CT u2 = u[0] * u[0] + u[1] * u[1];
    if (L::D == 3) {
        eu += CT(L::ez(q)) * u[2];
        u2 += u[2] * u[2];
    }
    return CT(L::w(q)) * (CT(3) * eu + CT(4.5) * eu * eu - CT(1.5) * u2);
}

// gneq <- M^-1 S M gneq, stress moments scaled by s (NsAllen.cu:416-431),
// evaluated as gneq + (s - 1) M^-1[:, stress] (M gneq)[stress] with integer
// rows of M and M^-1 numerators over a common denominator; the oracle uses
// the same form and operation order.
template <class L, class CT>
__device__ __forceinline__ void wmrt_filter(CT v[L::Q], CT s) {
    CT m[L::NSTRESS];
#pragma unroll
    for (int k = 0; k < L::NSTRESS; ++k) {
        CT acc = CT(0);
#pragma unroll
        for (int q = 0; q < L::Q; ++q) {
            const int c = L::M(L::stress(k), q);
            if (c == 1) acc += v[q];
            else if (c == -1) acc -= v[q];
            else if (c != 0) acc += CT(c) * v[q];
        }
        m[k] = acc;
    }
    const CT sm1 = s - CT(1);
#pragma unroll
    for (int q = 0; q < L::Q; ++q) {
        CT acc = CT(0);
#pragma unroll
        for (int k = 0; k < L::NSTRESS; ++k) {
            const int c = L::minv_num(q, L::stress(k));
            if (c == 1) acc += m[k];
            else if (c == -1) acc -= m[k];
            else if (c != 0) acc += CT(c) * m[k];
        }
        v[q] = v[q] + sm1 * cdiv<CT>(acc, double(L::minv_den(q)));
    }
}

// Viscous force of calculateVF (NsAllen.cu:102-132).
template <class L, class CT>
__device__ __forceinline__ void viscous_force(const CT gneq[L::Q], CT tau, CT drho, const CT dc[3], CT fv[3]) {
    CT sxx = 0, sxy = 0, syy = 0, szx = 0, szy = 0, szz = 0;
#pragma unroll
    for (int q = 0; q < L::Q; ++q) {
        const int x = L::ex(q), y = L::ey(q), z = L::ez(q);
        if (x * x) sxx += gneq[q];
        if (x * y) sxy += CT(x * y) * gneq[q];
        if (y * y) syy += gneq[q];
        if (L::D == 3) {
            if (z * y) szy += CT(z * y) * gneq[q];
            if (z * x) szx += CT(z * x) * gneq[q];
            if (z * z) szz += gneq[q];
        }
    }
    fv[0] = -tau * (sxx * dc[0] + sxy * dc[1] + szx * dc[2]) * drho;
    fv[1] = -tau * (sxy * dc[0] + syy * dc[1] + szy * dc[2]) * drho;
    fv[2] = (L::D == 3) ? -tau * (szz * dc[2] + szx * dc[0] + szy * dc[1]) * drho : CT(0);
}

template <class CT>
__device__ __forceinline__ CT tau_of(const Phys<CT>& ph, CT phi) {
    if (phi < CT(0)) return ph.tau_l;
    if (phi > CT(1)) return ph.tau_h;
    return ph.tau_l + phi * (ph.tau_h - ph.tau_l);
}

// Body force by case (NsAllen.cu:460-474); jet and unknown cases: 0.
template <class CT>
__device__ __forceinline__ void body_force(const Phys<CT>& ph, CT rho, CT fg[3]) {
    CT base;
    switch (ph.case_type) {
        case CLIP_CASE_RTI: base = rho; break;
        case CLIP_CASE_BUBBLE: base = rho - ph.rho_h; break;
        case CLIP_CASE_DROP: base = rho - ph.rho_l; break;
        default: base = CT(0); break;
    }
    fg[0] = base * ph.grav[0];
    fg[1] = base * ph.grav[1];
    fg[2] = base * ph.grav[2];
}

// phi stencil at interior node n: isotropic gradient (NsAllen.cu:280-322) and
// Laplacian of the chemical potential (NsAllen.cu:207-236); sums accumulate
// in lexicographic (x offset, y offset, z offset) order.
template <class L, class CT>
__device__ __forceinline__ void phi_stencil(const CT* __restrict__ phi, const DevGrid& g, long long n, CT& c0,
                                            CT dc[3], CT& lap, CT* nbq = nullptr) {
    c0 = phi[n];
    if (L::D == 2) {
        const long long sy = g.sy;
        const CT xp = phi[n + 1], xm = phi[n - 1], yp = phi[n + sy], ym = phi[n - sy];
        const CT pp = phi[n + 1 + sy], pm = phi[n + 1 - sy], mp = phi[n - 1 + sy], mm = phi[n - 1 - sy];
        if (nbq) {
            const CT v[3][3] = {{mm, xm, mp}, {ym, c0, yp}, {pm, xp, pp}};
#pragma unroll
            for (int q = 0; q < L::Q; ++q) nbq[q] = v[1 + L::ex(q)][1 + L::ey(q)];
        }
        dc[0] = cdiv<CT>(xp - xm, 3.0) + cdiv<CT>((pm + pp) - (mm + mp), 12.0);
        dc[1] = cdiv<CT>(yp - ym, 3.0) + cdiv<CT>((mp + pp) - (mm + pm), 12.0);
        dc[2] = CT(0);
        const CT corners = ((mm + mp) + pm) + pp;
        const CT edges = ((xm + ym) + yp) + xp;
        lap = cdiv<CT>(corners + CT(4) * edges - CT(20) * c0, 6.0);
    } else {
        const long long sy = g.sy, sz = g.sz;
        CT v[3][3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b)
#pragma unroll
                for (int d = 0; d < 3; ++d) v[a][b][d] = phi[n + (a - 1) + (b - 1) * sy + (d - 1) * sz];
        if (nbq) {
#pragma unroll
            for (int q = 0; q < L::Q; ++q) nbq[q] = v[1 + L::ex(q)][1 + L::ey(q)][1 + L::ez(q)];
        }
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            CT plus[3] = {0, 0, 0}, minus[3] = {0, 0, 0};
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b)
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        const int off[3] = {a - 1, b - 1, d - 1};
                        if (off[ax] == 0) continue;
                        const int others = (off[0] != 0) + (off[1] != 0) + (off[2] != 0) - 1;
                        if (off[ax] > 0) plus[others] += v[a][b][d];
                        else minus[others] += v[a][b][d];
                    }
            dc[ax] = CT(0.5) * (CT(4.0 / 9.0) * (plus[0] - minus[0]) + CT(1.0 / 9.0) * (plus[1] - minus[1]) +
                                CT(1.0 / 36.0) * (plus[2] - minus[2]));
        }
        CT faces = 0, edges = 0, corners = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b)
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const int m = (a != 1) + (b != 1) + (d != 1);
                    if (m == 1) faces += v[a][b][d];
                    else if (m == 2) edges += v[a][b][d];
                    else if (m == 3) corners += v[a][b][d];
                }
        lap = cdiv<CT>(CT(20) * faces + CT(6) * edges + corners - CT(200) * c0, 48.0);
    }
}

// ------------------------------------------------------------ node selection
// Threads per block of the step kernels; the bulk launches shape them as
// bx x by with bx = 256 where rows allow it, so that each block reads whole
// rows and a shifted pull straddles one extra 128-byte line per row only.
constexpr int BX = 64, BY = 4;

// Thread -> node of the bulk region: the interior box shrunk by one node
// along every axis that is not periodic on both sides (inset = 1), or the
// whole interior (inset = 0, used by the non-pulling kernels).
template <class L>
__device__ __forceinline__ bool node_of_thread(const DevGrid& g, const DevBC& bc, int inset, int c[3]) {
    const int ix = inset && !bc.both_periodic[0], iy = inset && !bc.both_periodic[1];
    const int iz = (L::D == 3) ? (inset && !bc.both_periodic[2]) : 0;
    c[0] = blockIdx.x * blockDim.x + threadIdx.x + g.dmin[0] + ix;
    c[1] = blockIdx.y * blockDim.y + threadIdx.y + g.dmin[1] + iy;
    c[2] = blockIdx.z + g.dmin[2] + iz;
    return c[0] <= g.dmax[0] - ix && c[1] <= g.dmax[1] - iy;
}

template <class L>
__host__ __device__ inline ShellDesc make_shell(const DevGrid& g, const DevBC& bc) {
    ShellDesc sd{};
    long long n[3];
    for (int a = 0; a < 3; ++a) n[a] = g.dmax[a] - g.dmin[a] + 1;
    for (int a = 0; a < L::D; ++a) {
        if (bc.both_periodic[a]) continue;
        long long cnt = 2;
        for (int b = 0; b < 3; ++b) {
            if (b == a) continue;
            bool earlier = false;
            for (int k = 0; k < sd.na; ++k) earlier |= sd.axes[k] == b;
            cnt *= earlier ? (n[b] - 2) : n[b];
        }
        sd.axes[sd.na] = a;
        sd.count[sd.na] = cnt;
        sd.na++;
    }
    return sd;
}

template <class L>
__device__ __forceinline__ bool shell_node_at(const DevGrid& g, const ShellDesc& sd, long long t64, int c[3]) {
    // 32-bit index arithmetic: a boundary layer holds far fewer than 2^32
    // nodes.  Every loop has constant bounds and every array index is a
    // constant, so ShellDesc and c stay in registers.
    unsigned t = (unsigned)t64;
    bool found = false;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (found || k >= sd.na) continue;
        const unsigned cnt = (unsigned)sd.count[k];
        if (t >= cnt) {
            t -= cnt;
            continue;
        }
        found = true;
        const int a = sd.axes[k];
        const unsigned half = cnt / 2;
        const bool hi = t >= half;
        unsigned r = hi ? t - half : t;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            if (b == a) continue;
            bool earlier = false;
#pragma unroll
            for (int m = 0; m < k; ++m) earlier |= sd.axes[m] == b;
            const int lo = g.dmin[b] + (earlier ? 1 : 0);
            const unsigned size = (unsigned)(g.dmax[b] - g.dmin[b] + 1 - (earlier ? 2 : 0));
            c[b] = lo + (int)(r % size);
            r /= size;
        }
        // arithmetic rather than a choice between two fields: a selected
        // field address would force the kernel arguments into scratch
#pragma unroll
        for (int b = 0; b < 3; ++b)
            if (b == a) c[b] = g.dmin[b] + (hi ? g.dmax[b] - g.dmin[b] : 0);
    }
    return found;
}

template <class L>
__device__ __forceinline__ bool shell_node(const DevGrid& g, const ShellDesc& sd, int c[3]) {
    return shell_node_at<L>(g, sd, (long long)blockIdx.x * blockDim.x + threadIdx.x, c);
}

// Source index of population q pulled by a bulk node: n - e_q, wrapped along
// periodic axes (kernelPeriodicBoundary + kernelStreaming, Solver.cu:75-141,
// NsAllen.cu:327-364).
struct PullOffsets {
    long long m[3], p[3];  // offset for e = +1 (source at -1) and e = -1 (source at +1)
};
template <class L>
__device__ __forceinline__ PullOffsets pull_offsets(const DevGrid& g, const int c[3]) {
    PullOffsets o;
    const long long st[3] = {1, g.sy, g.sz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const long long span = (long long)(g.dmax[a] - g.dmin[a]) * st[a];
        o.m[a] = (c[a] == g.dmin[a]) ? span : -st[a];
        o.p[a] = (c[a] == g.dmax[a]) ? -span : st[a];
    }
    return o;
}
template <class L>
__device__ __forceinline__ long long pull_src(long long n, const PullOffsets& o, int q) {
    long long s = n;
    if (L::ex(q) > 0) s += o.m[0];
    if (L::ex(q) < 0) s += o.p[0];
    if (L::ey(q) > 0) s += o.m[1];
    if (L::ey(q) < 0) s += o.p[1];
    if (L::ez(q) > 0) s += o.m[2];
    if (L::ez(q) < 0) s += o.p[2];
    return s;
}

// Bulk addressing of whole-row blocks (blockDim.y == 1): the row, the plane
// and their periodic wraps are the same for every lane of a block, only x
// varies.  Population addresses then split into a wave-uniform base (q plane,
// row, y/z offset) and a 32-bit lane offset shared by all q with the same e_x.
struct RowPull {
    long long row;              // c1 sy + c2 sz
    long long ym, yp, zm, zp;   // source offsets for e_y = +1 / -1, e_z = +1 / -1
    unsigned x, xm, xp;         // source x for e_x = 0, +1, -1
};
template <class L>
__device__ __forceinline__ RowPull row_pull(const DevGrid& g, const int c[3]) {
    RowPull r;
    r.row = c[1] * g.sy + c[2] * g.sz;
    const long long spy = (long long)(g.dmax[1] - g.dmin[1]) * g.sy;
    const long long spz = (long long)(g.dmax[2] - g.dmin[2]) * g.sz;
    r.ym = (c[1] == g.dmin[1]) ? spy : -g.sy;
    r.yp = (c[1] == g.dmax[1]) ? -spy : g.sy;
    r.zm = (L::D == 3 && c[2] == g.dmin[2]) ? spz : -g.sz;
    r.zp = (L::D == 3 && c[2] == g.dmax[2]) ? -spz : g.sz;
    r.x = (unsigned)c[0];
    r.xm = (unsigned)((c[0] == g.dmin[0]) ? g.dmax[0] : c[0] - 1);
    r.xp = (unsigned)((c[0] == g.dmax[0]) ? g.dmin[0] : c[0] + 1);
    return r;
}
template <class L>
__device__ __forceinline__ long long row_uniform(const RowPull& r, int q) {
    long long u = r.row;
    if (L::ey(q) > 0) u += r.ym;
    if (L::ey(q) < 0) u += r.yp;
    if (L::ez(q) > 0) u += r.zm;
    if (L::ez(q) < 0) u += r.zp;
    return u;
}
template <class L>
__device__ __forceinline__ unsigned row_lane(const RowPull& r, int q) {
    return L::ex(q) > 0 ? r.xm : (L::ex(q) < 0 ? r.xp : r.x);
}
// Element at (uniform base) + (32-bit lane index).
template <class T>
__device__ __forceinline__ T* at_lane(T* base, unsigned idx) {
    using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
    return (T*)((B*)base + idx * (unsigned)sizeof(T));
}

template <class L>
__device__ __forceinline__ bool node_of_thread_rows(const DevGrid& g, const DevBC& bc, int c[3]) {
    const int ix = !bc.both_periodic[0], iy = !bc.both_periodic[1];
    const int iz = (L::D == 3) ? !bc.both_periodic[2] : 0;
    c[0] = blockIdx.x * blockDim.x + threadIdx.x + g.dmin[0] + ix;
    c[1] = blockIdx.y + g.dmin[1] + iy;
    c[2] = blockIdx.z + g.dmin[2] + iz;
    return c[0] <= g.dmax[0] - ix;
}

// Node selection of the step kernels: SHELL = boundary layer (resolved pulls),
// !SHELL && !DIRECT = bulk (offset pulls, periodic wrap), DIRECT = every
// interior node (values read in place).
template <class L, bool DIRECT, bool SHELL, bool ROWS = false>
__device__ __forceinline__ bool select_node(const DevGrid& g, const DevBC& bc, const ShellDesc& sd, int c[3]) {
    if (SHELL) return shell_node<L>(g, sd, c);
    if (ROWS) return node_of_thread_rows<L>(g, bc, c);
    return node_of_thread<L>(g, bc, DIRECT ? 0 : 1, c);
}

// ------------------------------------------------------------ step kernels
// The phi ghost refresh of a step without inlet sides (one k_ghost_resolve
// pass, periodic and mirror copies from the old phi) carried by extra blocks
// of the phase launch: it writes only the new phi's ghost ring and reads only
// the old phi, while the phase sum writes the new phi's interior from g, so
// one launch does both.  Blocks from `first` on (the z index of the box grid
// in k_phase, the block index in k_phase_pair) visit the ghost cells.
template <class CT>
__device__ inline void ghost_resolve_cell(const CT* oldp, CT* newp, const DevGrid& g, const DevBC& bc, int D,
                                          long long t, int do_periodic, int do_mirror, int ghost_from_new);
template <class CT>
struct GhostJob {
    const CT* oldp;
    CT* newp;
    long long total;  // ghost cells; 0: no ghost blocks in this launch
    unsigned first;
};
template <class CT>
__device__ __forceinline__ void ghost_job_block(const GhostJob<CT>& gj, const DevGrid& g, const DevBC& bc, int D,
                                                long long block) {
    const long long t = block * (blockDim.x * blockDim.y) + threadIdx.y * blockDim.x + threadIdx.x;
    if (t < gj.total) ghost_resolve_cell(gj.oldp, gj.newp, g, bc, D, t, 1, 1, 0);
}

// phi^{n+1} = sum_q g_post (kernelMacroscopicg, NsAllen.cu:516-540).
template <class L, class ST, class CT, class SP, bool DIRECT, bool SHELL, bool ROWS = false>
__global__ __launch_bounds__(BX* BY) void k_phase(StepArgs<ST, CT> a, GhostJob<CT> gj) {
    if (gj.total && blockIdx.z >= gj.first) {
        ghost_job_block(gj, a.g, a.bc, L::D,
                        ((long long)(blockIdx.z - gj.first) * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x);
        return;
    }
    int c[3];
    if (!select_node<L, DIRECT, SHELL, ROWS>(a.g, a.bc, a.shell, c)) return;
    const long long n = a.g.idx(c[0], c[1], c[2]);
    CT sum = CT(0);
    if (DIRECT) {
#pragma unroll
        for (int q = 0; q < L::Q; ++q)
            sum += SP::ld(a.gin, (long long)q * a.g.nodes + n, q, gref<SP>(a.phi_ref, ref_index<L>(a.g, a.bc, c, q)));
    } else if (!SHELL && ROWS) {
        const RowPull rp = row_pull<L>(a.g, c);
        const CT r = gref<SP>(a.phi_ref, n);  // every pulled g is referenced to this node
#pragma unroll
        for (int q = 0; q < L::Q; ++q)
            sum += SP::ld(at_lane(a.gin + (long long)q * a.g.nodes + row_uniform<L>(rp, q), row_lane<L>(rp, q)), 0, q, r);
    } else if (!SHELL) {
        const PullOffsets po = pull_offsets<L>(a.g, c);
        const CT r = gref<SP>(a.phi_ref, n);  // every pulled g is referenced to this node
#pragma unroll
        for (int q = 0; q < L::Q; ++q) {
            const long long src = pull_src<L>(n, po, q);
            sum += SP::ld(a.gin, (long long)q * a.g.nodes + src, q, r);
        }
    } else {
        Puller<L, ST, CT, SP, true> P{a.gin, a.g, a.bc, a.ufc, a.prev_g_in, a.phi_ref};
#pragma unroll
        for (int q = 0; q < L::Q; ++q) sum += P.lvl3(c, q);
    }
    a.phi[n] = sum;
}

// Boundary layer, first half of the step: post-streaming f and g of every
// (shell node, q) pair with the boundary fix-ups (Puller), into a dense
// scratch that the shell launch of k_fused reads, and the outflow planes of
// the next step.  One thread per pair spreads the fix-up chains over the
// chip; keeping them out of k_fused keeps that kernel within its registers.
// blockIdx.y is q, so every wave resolves one compile-time direction.
template <class L, class ST, class CT, class SP, int Q0>
__device__ __forceinline__ void shell_pull_dir(const ST* fin, const ST* gin, const DevGrid& g, const DevBC& bc,
                                               const CT* uin, const CT* const* prev_f_in,
                                               const CT* const* prev_g_in, const CT* phi_ref, CT* shell_f,
                                               CT* shell_g, CT* const* prev_f_out, CT* const* prev_g_out,
                                               unsigned t, unsigned S, const int c[3]) {
    Puller<L, ST, CT, SP, false> Pf{fin, g, bc, uin, prev_f_in, phi_ref};
    Puller<L, ST, CT, SP, true> Pg{gin, g, bc, uin, prev_g_in, phi_ref};
    const CT fv = Pf.lvl3(c, Q0);
    const CT gv = Pg.lvl3(c, Q0);
    shell_f[(long long)Q0 * S + t] = fv;
    shell_g[(long long)Q0 * S + t] = gv;
    if (bc.is_fc) {
        store_prev<L, CT>(g, bc, prev_f_out, c, Q0, fv);
        store_prev<L, CT>(g, bc, prev_g_out, c, Q0, gv);
    }
}
template <class L, class ST, class CT, class SP>
__global__ __launch_bounds__(BX* BY) void k_shell_pull(StepArgs<ST, CT> a) {
    const unsigned S = (unsigned)shell_total(a.shell);
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    int c[3];
    if (t >= S || !shell_node_at<L>(a.g, a.shell, t, c)) return;
    // outflow-plane pointers as local arrays (constant indices keep them in registers)
    const CT* pfi[6];
    const CT* pgi[6];
    CT* pfo[6];
    CT* pgo[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        pfi[s] = a.prev_f_in[s];
        pgi[s] = a.prev_g_in[s];
        pfo[s] = a.prev_f_out[s];
        pgo[s] = a.prev_g_out[s];
    }
#define CLIP_SHELL_DIR(k)                                                                                   \
    case k:                                                                                                 \
        if constexpr (k < L::Q)                                                                             \
            shell_pull_dir<L, ST, CT, SP, k>(a.fin, a.gin, a.g, a.bc, a.ufc, pfi, pgi, a.phi_ref, a.shell_f, \
                                             a.shell_g, pfo, pgo, t, S, c);                                 \
        break;
    switch (blockIdx.y) {
        CLIP_SHELL_DIR(0) CLIP_SHELL_DIR(1) CLIP_SHELL_DIR(2) CLIP_SHELL_DIR(3) CLIP_SHELL_DIR(4)
        CLIP_SHELL_DIR(5) CLIP_SHELL_DIR(6) CLIP_SHELL_DIR(7) CLIP_SHELL_DIR(8) CLIP_SHELL_DIR(9)
        CLIP_SHELL_DIR(10) CLIP_SHELL_DIR(11) CLIP_SHELL_DIR(12) CLIP_SHELL_DIR(13) CLIP_SHELL_DIR(14)
        CLIP_SHELL_DIR(15) CLIP_SHELL_DIR(16) CLIP_SHELL_DIR(17) CLIP_SHELL_DIR(18)
        default: break;
    }
#undef CLIP_SHELL_DIR
}

// phi^{n+1} = sum_q g_post at the boundary-layer nodes (q ascending, as
// k_phase sums).
template <class L, class ST, class CT>
__global__ __launch_bounds__(BX* BY) void k_shell_phi(StepArgs<ST, CT> a) {
    const long long S = shell_total(a.shell);
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int c[3];
    if (t >= S || !shell_node_at<L>(a.g, a.shell, t, c)) return;
    CT sum = CT(0);
#pragma unroll
    for (int q = 0; q < L::Q; ++q) sum += a.shell_g[(long long)q * S + t];
    a.phi[a.g.idx(c[0], c[1], c[2])] = sum;
}

// Reference phase field of each stored g_q of node c (mixed storage): phi at
// the node that pulls it next, c + e_q, wrapped along the axes periodic on
// both ends.  nbq (phi at n + e_q, the stencil's own loads) is that value
// away from those seams; at a seam the wraps are per-axis deltas.
template <class L, class ST, class CT>
__device__ __forceinline__ void g_store_refs(const StepArgs<ST, CT>& a, const int c[3], long long n, const CT nbq[],
                                             CT refq[]) {
    constexpr int Q = L::Q;
    const DevGrid& g = a.g;
    bool seam = false;
#pragma unroll
    for (int ax = 0; ax < L::D; ++ax)
        seam |= a.bc.both_periodic[ax] && (c[ax] == g.dmin[ax] || c[ax] == g.dmax[ax]);
    if (!seam) {
#pragma unroll
        for (int q = 0; q < Q; ++q) refq[q] = nbq[q];
        return;
    }
    long long dp[3], dm[3];  // delta for e_a = +1 / -1
    const long long st3[3] = {1, g.sy, g.sz};
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const bool bp = ax < L::D && a.bc.both_periodic[ax];
        const long long span = (long long)(g.dmax[ax] - g.dmin[ax] + 1) * st3[ax];
        dp[ax] = (bp && c[ax] == g.dmax[ax]) ? -span : 0;
        dm[ax] = (bp && c[ax] == g.dmin[ax]) ? span : 0;
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        long long r = n + L::ex(q) + L::ey(q) * g.sy + L::ez(q) * g.sz;
#pragma unroll
        for (int ax = 0; ax < L::D; ++ax) {
            const int e = L::e(q, ax);
            if (e > 0) r += dp[ax];
            if (e < 0) r += dm[ax];
        }
        refq[q] = a.phi[r];
    }
}

enum : int { MODE_FULL = 0, MODE_COLLIDE = 1 };

// Fused macroscopic + collision (kernelMacroscopicf NsAllen.cu:369-511,
// normal_FD 241-275, kernelCollisionMRTg 545-584, kernelCollisionMRTf 589-763).
// FP64: the reference's transforms and operation order.  FP32 / mixed: the
// exact moment-space form of the same step (node.hpp).
// The node of this thread and its update.  CARRIED: a boundary-layer node
// taken by the extra z-slices (from shell_first on) of a bulk launch.  TILE:
// the phi stencil reads `tile`, the block's phi^{n+1} with a one-node halo in
// LDS, BX + 2 values a row (k_step_tile, BX x BY blocks), instead of the phi
// array, and
// the pulled f and g are the caller's registers pre_f / pre_g.
template <class L, class ST, class CT, class SP, bool DIRECT, int MODE, bool SHELL, bool ROWS, bool CARRIED,
          bool TILE = false>
__device__ __forceinline__ void fused_node(const StepArgs<ST, CT>& a, unsigned shell_first,
                                           const CT* tile = nullptr, const CT* pre_f = nullptr,
                                           const CT* pre_g = nullptr) {
    constexpr int Q = L::Q;
    constexpr bool MOMENT_FORM = sizeof(CT) == 4;
    int c[3];
    long long shell_t = 0;
    if constexpr (CARRIED) {
        const long long block =
            ((long long)(blockIdx.z - shell_first) * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        shell_t = block * (blockDim.x * blockDim.y) + threadIdx.y * blockDim.x + threadIdx.x;
        if (!shell_node_at<L>(a.g, a.shell, shell_t, c)) return;
    } else if constexpr (TILE) {
        if (!node_of_thread<L>(a.g, a.bc, 0, c)) return;  // the tile spans the whole interior
    } else {
        if (!select_node<L, DIRECT, SHELL, ROWS>(a.g, a.bc, a.shell, c)) return;
        if (SHELL) shell_t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    }
    const DevGrid& g = a.g;
    const Phys<CT>& ph = a.ph;
    const long long n = g.idx(c[0], c[1], c[2]);
    // outflow planes are written only by boundary-layer nodes
    const bool edge = SHELL || (DIRECT && on_edge<L>(g, c));

    // phi stencil -> grad phi, chemical potential
    CT c0, dc[3], lap;
    CT nbq[Q];  // phi at n + e_q (the stencil's own loads)
    if constexpr (TILE) {
        static_assert(L::D == 2 && !SP::uses_ref, "the LDS tile serves the 2D plain-storage step");
        DevGrid tg = g;
        tg.sy = BX + 2;
        phi_stencil<L, CT>(tile, tg, (long long)(threadIdx.y + 1) * tg.sy + threadIdx.x + 1, c0, dc, lap, nullptr);
    } else {
        phi_stencil<L, CT>(a.phi, g, n, c0, dc, lap, SP::uses_ref ? nbq : nullptr);
    }
    // reference of the stored g (mixed policy): phi of the node that pulls it
    CT refq[SP::uses_ref ? Q : 1];
    if constexpr (SP::uses_ref) g_store_refs<L>(a, c, n, nbq, refq);
    auto gout_ref = [&](int q) -> CT {
        if constexpr (!SP::uses_ref) return CT(0);
        else return refq[q];
    };

    PullOffsets po;
    if (!DIRECT && !SHELL && !ROWS) po = pull_offsets<L>(g, c);
    RowPull rp;
    if (ROWS) rp = row_pull<L>(g, c);
    // population q of this node (stores) and of its pull source (bulk loads)
    auto dst_ptr = [&](ST* A, int q) -> ST* {
        if constexpr (ROWS) return at_lane(A + (long long)q * g.nodes + rp.row, rp.x);
        else return A + (long long)q * g.nodes + n;
    };
    auto src_ptr = [&](const ST* A, int q) -> const ST* {
        if constexpr (ROWS) return at_lane(A + (long long)q * g.nodes + row_uniform<L>(rp, q), row_lane<L>(rp, q));
        else return A + (long long)q * g.nodes + pull_src<L>(n, po, q);
    };

    // f_post
    CT fp[Q];
    {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            if (DIRECT) fp[q] = SP::ld(a.fin, (long long)q * g.nodes + n, q, CT(0));
            else if (!SHELL)
                fp[q] = TILE ? pre_f[q] : SP::ld(src_ptr(a.fin, q), 0, q, CT(0));
            else
                fp[q] = a.shell_f[(long long)q * shell_total(a.shell) + shell_t];
        }
        if (DIRECT && a.bc.is_fc && edge) {
#pragma unroll
            for (int q = 0; q < Q; ++q) store_prev<L, CT>(g, a.bc, a.prev_f_out, c, q, fp[q]);
        }
    }
    // g_post
    const CT rn = (!DIRECT && !SHELL) ? gref<SP>(a.phi_ref, n) : CT(0);
    auto pull_g = [&](int q) -> CT {
        if (DIRECT) return SP::ld(a.gin, (long long)q * g.nodes + n, q, gref<SP>(a.phi_ref, ref_index<L>(g, a.bc, c, q)));
        if (TILE) return pre_g[q];
        if (!SHELL) return SP::ld(src_ptr(a.gin, q), 0, q, rn);
        return a.shell_g[(long long)q * shell_total(a.shell) + shell_t];
    };

    if constexpr (MOMENT_FORM) {
        CT p;
        if (MODE == MODE_FULL) {
            double pd = 0.0;  // pressure recovered in FP64
#pragma unroll
            for (int q = 0; q < Q; ++q) pd += (double)fp[q];
            p = (CT)pd;
        } else {
            p = a.p[n];
        }
        CT mom[3], Pf[3][3];
        pop_moments<L>(fp, mom, Pf);
        const NodeScalars<CT> z = node_scalars<L>(ph, c0, lap);
        CT u[3];
        if (MODE == MODE_FULL) {
            CT uo[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) uo[d] = (d < L::D) ? a.uin[(long long)d * g.nodes + n] : CT(0);
            velocity_of<L>(z, p, mom, Pf, dc, uo, u);
#pragma unroll
            for (int d = 0; d < L::D; ++d) a.uout[(long long)d * g.nodes + n] = u[d];
        } else {
#pragma unroll
            for (int d = 0; d < 3; ++d) u[d] = (d < L::D) ? a.ucol[(long long)d * g.nodes + n] : CT(0);
        }
        CT gp[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            gp[q] = pull_g(q);
            if (DIRECT && a.bc.is_fc && edge) store_prev<L, CT>(g, a.bc, a.prev_g_out, c, q, gp[q]);
        }
        collide_node<L, CT>(
            z, p, Pf, c0, dc, u, gp,
            [&](int q, CT v) { SP::st(dst_ptr(a.gout, q), 0, q, v, gout_ref(q)); },
            [&](int q, CT v) { SP::st(dst_ptr(a.fout, q), 0, q, v, CT(0)); });
        return;
    } else {
        const CT mu = CT(4) * ph.beta * c0 * (c0 - CT(1)) * (c0 - CT(0.5)) - ph.kappa * lap;
        const CT rho = ph.rho_l + c0 * (ph.rho_h - ph.rho_l);
        const CT drho = ph.rho_h - ph.rho_l;
        const CT drho3 = cdiv<CT>(drho, 3.0);
        const CT rr = CT(1) / rho;
        const CT tau = tau_of(ph, c0);
        const CT s = CT(1) / (tau + CT(0.5));
        CT fg[3];
        body_force(ph, rho, fg);

        CT p, u[3];
        if (MODE == MODE_FULL) {
            // macroscopic p, u (kernelMacroscopicf): gneq from u_old
            p = CT(0);
#pragma unroll
            for (int q = 0; q < Q; ++q) p += fp[q];
            CT uo[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) uo[d] = (d < L::D) ? a.uin[(long long)d * g.nodes + n] : CT(0);
            CT gneq[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) gneq[q] = fp[q] - (p * CT(L::w(q)) + gamma_u<L, CT>(q, uo));
            CT fv[3];
            wmrt_filter<L, CT>(gneq, s);
            viscous_force<L, CT>(gneq, tau, drho, dc, fv);
            CT F[3], mom[3] = {0, 0, 0};
#pragma unroll
            for (int d = 0; d < 3; ++d) F[d] = mu * dc[d] + (-p * drho3 * dc[d]) + fv[d] + fg[d];
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                if (L::ex(q)) mom[0] += CT(L::ex(q)) * fp[q];
                if (L::ey(q)) mom[1] += CT(L::ey(q)) * fp[q];
                if (L::ez(q)) mom[2] += CT(L::ez(q)) * fp[q];
            }
#pragma unroll
            for (int d = 0; d < 3; ++d) u[d] = (d < L::D) ? mom[d] + div_rn(CT(0.5) * F[d], rho, rr) : CT(0);
#pragma unroll
            for (int d = 0; d < L::D; ++d) a.uout[(long long)d * g.nodes + n] = u[d];
        } else {
            p = a.p[n];
#pragma unroll
            for (int d = 0; d < 3; ++d) u[d] = (d < L::D) ? a.ucol[(long long)d * g.nodes + n] : CT(0);
        }

        // normal (normal_FD)
        CT nr[3];
        {
            CT sq = dc[0] * dc[0] + dc[1] * dc[1];
            if (L::D == 3) sq += dc[2] * dc[2];
            const CT t = sqrt(sq) + CT(1e-32);
            const CT rt = CT(1) / t;
            nr[0] = div_rn(dc[0], t, rt);
            nr[1] = div_rn(dc[1], t, rt);
            nr[2] = (L::D == 3) ? div_rn(dc[2], t, rt) : CT(0);
        }

        // g collision (kernelCollisionMRTg), one population at a time
        {
            const CT wc = CT(1) / (CT(0.5) + CT(3) * ph.mobility);
            const CT aterm = (CT(3) * ph.mobility * (CT(1) - CT(4) * ((c0 - CT(0.5)) * (c0 - CT(0.5))))) / ph.W;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const CT gp = pull_g(q);
                if (DIRECT && a.bc.is_fc && edge) store_prev<L, CT>(g, a.bc, a.prev_g_out, c, q, gp);
                CT en = CT(L::ex(q)) * nr[0] + CT(L::ey(q)) * nr[1];
                if (L::D == 3) en += CT(L::ez(q)) * nr[2];
                const CT heq = c0 * (gamma_u<L, CT>(q, u) + CT(L::w(q))) + CT(L::w(q)) * (aterm * en);
                SP::st(dst_ptr(a.gout, q), 0, q, gp * (CT(1) - wc) + heq * wc, CT(0));
            }
        }

        // f collision (kernelCollisionMRTf)
        CT ga[Q], gneq[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            ga[q] = gamma_u<L, CT>(q, u);
            gneq[q] = fp[q] - (p * CT(L::w(q)) + ga[q]);
        }
        wmrt_filter<L, CT>(gneq, s);
        CT fv[3];
        viscous_force<L, CT>(gneq, tau, drho, dc, fv);
        CT F[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) F[d] = mu * dc[d] + (-p * drho3 * dc[d]) + fv[d] + fg[d];
        CT hlp[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            CT eF = CT(L::ex(q)) * F[0] + CT(L::ey(q)) * F[1];
            if (L::D == 3) eF += CT(L::ez(q)) * F[2];
            hlp[q] = div_rn(CT(3) * CT(L::w(q)) * eF, rho, rr);
            gneq[q] = fp[q] - (p * CT(L::w(q)) + ga[q] - CT(0.5) * hlp[q]);
        }
        wmrt_filter<L, CT>(gneq, s);
#pragma unroll
        for (int q = 0; q < Q; ++q) SP::st(dst_ptr(a.fout, q), 0, q, fp[q] - gneq[q] + hlp[q], CT(0));
    }
}

// The launch over the bulk box (or over the boundary layer, SHELL).  A bulk
// launch may carry the boundary layer too (CARRY): blocks from z index
// shell_first on take its nodes, which the bulk nodes neither read nor write
// in this kernel, so one launch collides both.  The carrying instantiation
// holds both code paths and the registers of the larger one.
template <class L, class ST, class CT, class SP, bool DIRECT, int MODE, bool SHELL, bool ROWS = false,
          bool CARRY = false>
__global__ __launch_bounds__(BX* BY) void k_fused(StepArgs<ST, CT> a, unsigned shell_first) {
    if constexpr (CARRY && !SHELL && !DIRECT) {
        if (blockIdx.z >= shell_first) {
            fused_node<L, ST, CT, SP, DIRECT, MODE, true, false, true>(a, shell_first);
            return;
        }
    }
    fused_node<L, ST, CT, SP, DIRECT, MODE, SHELL, ROWS, false>(a, 0u);
}

// ------------------------------------------------------------ node pairs
// 16-bit storage moves one population of one node per lane and access, and
// a 19-plane pull at that width is bound by the access count rather than by
// bytes.  k_phase_pair gives each lane two x-adjacent bulk nodes, so each
// pulled g of both is one 32-bit access, split into two 16-bit loads where a
// periodic wrap separates the sources.  The word is 2-byte aligned in
// general; gfx950 global memory serves such loads (measured ~3% slower than
// aligned ones, with identical bytes).  Per node it sums exactly as k_phase
// does.
struct PairGrid {
    int xlo, xhi;      // bulk x range
    int ylo, ny, zlo;  // bulk rows: y = ylo + r % ny, z = zlo + r / ny
    unsigned per_row;  // pairs per row: (xhi - xlo + 2) / 2
    unsigned total;    // pairs in the bulk
};

// Pair number `pair`: node 0 at c, node 1 at c + (1,0,0) (absent when the
// row has an odd length and this is its last pair).
__device__ __forceinline__ void pair_coords(const PairGrid& pg, unsigned pair, int c[3], bool& has1) {
    const unsigned r = pair / pg.per_row;
    const unsigned p = pair - r * pg.per_row;
    c[0] = pg.xlo + 2 * (int)p;
    c[1] = pg.ylo + (int)(r % (unsigned)pg.ny);
    c[2] = pg.zlo + (int)(r / (unsigned)pg.ny);
    has1 = c[0] + 1 <= pg.xhi;
}
__device__ __forceinline__ bool pair_of_thread(const PairGrid& pg, int c[3], bool& has1) {
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= pg.total) return false;
    pair_coords(pg, t, c, has1);
    return true;
}

// Pull sources of the pair along x for direction q (element index within the
// row): node 0's source, wrapped on x when x is periodic on both ends, and
// node 1's (node 0's + 1 when node 1 is absent, which stays inside the row).
template <class L>
__device__ __forceinline__ void pair_sources(const DevGrid& g, const int c[3], bool has1, int q, int& s0, int& s1) {
    const int ex = L::ex(q);
    s0 = c[0] - ex;
    s1 = c[0] + 1 - ex;
    // the bulk reaches dmin / dmax only along an x periodic on both ends
    if (ex > 0 && c[0] == g.dmin[0]) s0 = g.dmax[0];
    if (ex < 0 && c[0] == g.dmax[0]) s0 = g.dmin[0];  // node 0 last of an odd row
    if (ex < 0 && c[0] + 1 == g.dmax[0]) s1 = g.dmin[0];
    if (!has1) s1 = s0 + 1;
}

// Two 16-bit values at row + s0 (low half) and row + s1 (high half) of one plane.
template <class ST>
__device__ __forceinline__ unsigned load_pair_word(const ST* row, int s0, int s1) {
    unsigned w;
    if (s1 == s0 + 1) {
        __builtin_memcpy(&w, row + s0, 4);
    } else {
        unsigned short h0, h1;
        __builtin_memcpy(&h0, row + s0, 2);
        __builtin_memcpy(&h1, row + s1, 2);
        w = (unsigned)h0 | ((unsigned)h1 << 16);
    }
    return w;
}
template <class ST>
__device__ __forceinline__ ST half_of(unsigned w, int hi) {
    const unsigned short h = (unsigned short)(hi ? w >> 16 : w & 0xffffu);
    ST v;
    __builtin_memcpy(&v, &h, 2);
    return v;
}
template <class ST>
__device__ __forceinline__ void load_pair(const ST* row, int s0, int s1, ST& v0, ST& v1) {
    const unsigned w = load_pair_word(row, s0, s1);
    v0 = half_of<ST>(w, 0);
    v1 = half_of<ST>(w, 1);
}

// Source row (y, z part of the pull) of direction q, relative to the pair's row.
template <class L>
__device__ __forceinline__ long long pair_row_offset(const PullOffsets& po, int q) {
    long long s = 0;
    if (L::ey(q) > 0) s += po.m[1];
    if (L::ey(q) < 0) s += po.p[1];
    if (L::ez(q) > 0) s += po.m[2];
    if (L::ez(q) < 0) s += po.p[2];
    return s;
}

template <class L, class ST, class SP>
__global__ __launch_bounds__(BX* BY) void k_phase_pair(StepArgs<ST, float> a, PairGrid pg, GhostJob<float> gj) {
    if (gj.total && blockIdx.x >= gj.first) {
        ghost_job_block(gj, a.g, a.bc, L::D, (long long)(blockIdx.x - gj.first));
        return;
    }
    int c[3];
    bool has1;
    if (!pair_of_thread(pg, c, has1)) return;
    const DevGrid& g = a.g;
    const long long n0 = g.idx(c[0], c[1], c[2]);
    const long long row = n0 - c[0];
    const PullOffsets po = pull_offsets<L>(g, c);
    const float r0 = gref<SP>(a.phi_ref, n0), r1 = has1 ? gref<SP>(a.phi_ref, n0 + 1) : 0.0f;
    float sum0 = 0.0f, sum1 = 0.0f;
#pragma unroll
    for (int q = 0; q < L::Q; ++q) {
        int s0, s1;
        pair_sources<L>(g, c, has1, q, s0, s1);
        ST v0, v1;
        load_pair(a.gin + (long long)q * g.nodes + row + pair_row_offset<L>(po, q), s0, s1, v0, v1);
        sum0 += SP::ld(&v0, 0, q, r0);
        sum1 += SP::ld(&v1, 0, q, r1);
    }
    a.phi[n0] = sum0;
    if (has1) a.phi[n0 + 1] = sum1;
}

// ------------------------------------------------------------ phi ghost ring
// Visits every cell of the ghost planes of the first 2*D sides (edge cells
// more than once; every visit writes the same value).
__device__ inline bool ghost_cell_of_thread(const DevGrid& g, int D, long long t, int c[3]) {
    for (int side = 0; side < 2 * D; ++side) {
        const long long ps = plane_size(g, side);
        if (t < ps) {
            const int a = side >> 1;
            const int t1 = (a == 0) ? 1 : 0;
            const int t2 = (a == 2) ? 1 : 2;
            c[t1] = (int)(t % g.ext[t1]);
            c[t2] = (int)(t / g.ext[t1]);
            c[a] = (side & 1) ? g.dmax[a] + 1 : g.dmin[a] - 1;
            return true;
        }
        t -= ps;
    }
    return false;
}

// Reads phi at `src` for a ghost refresh: interior cells from the old phi,
// ghost cells from the old phi (values of the previous step) or, in the
// launches after the first, from the new phi (written earlier this step).
template <class CT>
__device__ inline CT ghost_source(const CT* oldp, const CT* newp, const DevGrid& g, int D, const int s[3],
                                  int ghost_from_new) {
    const long long i = g.idx(s[0], s[1], s[2]);
    return (!g.interior(s, D) && ghost_from_new) ? newp[i] : oldp[i];
}

// Periodic ghost copy (Solver.cu:75-141) and/or mirror copy
// (Solver.cu:927-978) of phi: ghost <- phi at the cell with every periodic
// axis wrapped and every mirror axis clamped.  The source cell is never
// written by the same launch.
// The cell whose phi ghost cell c copies: every periodic axis wrapped and
// every mirror axis clamped; false when neither applies (the cell keeps its
// value).
__device__ inline bool ghost_source_cell(const DevGrid& g, const DevBC& bc, int D, const int c[3], int do_periodic,
                                         int do_mirror, int s[3]) {
    s[0] = c[0];
    s[1] = c[1];
    s[2] = c[2];
    bool changed = false;
    for (int a = 0; a < D; ++a) {
        const bool lo = c[a] < g.dmin[a], hi = c[a] > g.dmax[a];
        if (!lo && !hi) continue;
        // the mirror copy runs after the periodic copy, so on a mirror-type
        // side it decides the value (NsAllen.cu:884-887)
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
    return changed;
}
template <class CT>
__device__ inline void ghost_resolve_cell(const CT* oldp, CT* newp, const DevGrid& g, const DevBC& bc, int D,
                                          long long t, int do_periodic, int do_mirror, int ghost_from_new) {
    int c[3];
    if (!ghost_cell_of_thread(g, D, t, c)) return;
    int s[3];
    if (ghost_source_cell(g, bc, D, c, do_periodic, do_mirror, s))
        newp[g.idx(c[0], c[1], c[2])] = ghost_source(oldp, newp, g, D, s, ghost_from_new);
}
template <class CT>
__global__ void k_ghost_resolve(const CT* oldp, CT* newp, DevGrid g, DevBC bc, int D, long long total,
                                int do_periodic, int do_mirror, int ghost_from_new) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < total) ghost_resolve_cell(oldp, newp, g, bc, D, t, do_periodic, do_mirror, ghost_from_new);
}

// One launch per step for a 2D domain periodic along x (plain storage, no
// inlet or free-convect side).  Each BX x BY block sums phi^{n+1} (k_phase's
// pull sum, same loads and order) over its tile and a one-node halo into LDS;
// halo cells in the ghost ring take what the ghost job gives them (phi^n at
// the wrapped or mirrored node, ghost_source_cell).  The block writes its
// tile's bulk nodes and the ring cells of its halo to the phi array (cells
// shared by several halos receive the same value from each), then collides
// its tile with the stencil read from LDS and, for each node, the pulled f
// and g that its own sum loaded.  LAYER: y is not periodic on both sides, so
// the rows next to those sides are the boundary layer, whose phi^{n+1} and
// post-streaming f and g k_shell_phi / k_shell_pull have already produced
// (phi array, layer scratch); those rows collide here too, from that scratch.
// Interior halo cells are summed again by every block whose halo holds them,
// with the same result, so the launch equals k_phase followed by k_fused.
// Reads: the pull buffers, the old phi and the layer scratch; writes: the new
// phi (bulk and ring) and the other population buffers, so blocks never read
// what another block writes.
template <class L, class ST, class CT, class SP, bool LAYER>
__global__ __launch_bounds__(BX* BY) void k_step_tile(StepArgs<ST, CT> a, const CT* phi_old) {
    static_assert(L::D == 2 && !SP::uses_ref, "the tiled step serves 2D plain storage");
    constexpr int TX = BX + 2, TY = BY + 2, HALO = 2 * TX + 2 * BY;
    static_assert(HALO <= BX * BY, "one halo cell per thread at most");
    __shared__ CT tile[TX * TY];
    const DevGrid& g = a.g;
    const int x0 = (int)(blockIdx.x * BX) + g.dmin[0] - 1, y0 = (int)(blockIdx.y * BY) + g.dmin[1] - 1;
    const int tid = threadIdx.y * BX + threadIdx.x;
    // cell 0: this thread's own node; cell 1: halo cell `tid` (top row, bottom
    // row, then the left and right columns).  Both load before either result
    // is stored.
    int hx[2], hy[2];
    hx[0] = threadIdx.x + 1;
    hy[0] = threadIdx.y + 1;
    if (tid < TX) {
        hx[1] = tid;
        hy[1] = 0;
    } else if (tid < 2 * TX) {
        hx[1] = tid - TX;
        hy[1] = TY - 1;
    } else {
        const int k = tid - 2 * TX;
        hx[1] = (k & 1) ? TX - 1 : 0;
        hy[1] = 1 + (k >> 1);
    }
    CT v[2];
    long long nn[2];
    bool use[2], store[2];
    // the pulled f and g of this thread's node (cell 0), which its collision uses
    CT pf[L::Q], pg[L::Q];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const int c[3] = {x0 + hx[e], y0 + hy[e], g.dmin[2]};
        // beyond the ring of a partial tile: no node the tile's stencils read
        use[e] = (e == 0 || tid < HALO) && c[0] <= g.dmax[0] + 1 && c[1] <= g.dmax[1] + 1;
        nn[e] = g.idx(c[0], c[1], c[2]);
        v[e] = CT(0);
        store[e] = false;
        if (!use[e]) continue;
        const bool layer = LAYER && (c[1] == g.dmin[1] || c[1] == g.dmax[1]);
        if (!g.interior(c, 2)) {
            int s[3] = {c[0], c[1], c[2]};
            bool copied = true;
            if constexpr (LAYER) {
                copied = ghost_source_cell(g, a.bc, 2, c, 1, 1, s);
            } else {
                // every side periodic: the copy is the wrap on both axes
#pragma unroll
                for (int ax = 0; ax < 2; ++ax)
                    s[ax] = c[ax] < g.dmin[ax] ? g.dmax[ax] : (c[ax] > g.dmax[ax] ? g.dmin[ax] : c[ax]);
            }
            if (copied) {
                v[e] = phi_old[g.idx(s[0], s[1], s[2])];
                store[e] = true;
            } else {
                v[e] = a.phi[nn[e]];  // a ring cell no copy reaches keeps its value
            }
        } else if (layer) {
            v[e] = a.phi[nn[e]];
            if (e == 0) {
                // the layer's scratch: the lower row, then the upper one
                const long long S = shell_total(a.shell), n0 = g.dmax[0] - g.dmin[0] + 1;
                const long long t = (c[1] == g.dmax[1] ? n0 : 0) + (c[0] - g.dmin[0]);
#pragma unroll
                for (int q = 0; q < L::Q; ++q) {
                    pf[q] = a.shell_f[(long long)q * S + t];
                    pg[q] = a.shell_g[(long long)q * S + t];
                }
            }
        } else {
            const PullOffsets po = pull_offsets<L>(g, c);
            const CT r = gref<SP>(a.phi_ref, nn[e]);
            if (e == 0) {
#pragma unroll
                for (int q = 0; q < L::Q; ++q) {
                    const long long src = (long long)q * g.nodes + pull_src<L>(nn[0], po, q);
                    pg[q] = SP::ld(a.gin, src, q, r);
                    pf[q] = SP::ld(a.fin, src, q, CT(0));
                }
#pragma unroll
                for (int q = 0; q < L::Q; ++q) v[0] += pg[q];
                store[0] = true;
            } else {
#pragma unroll
                for (int q = 0; q < L::Q; ++q)
                    v[1] += SP::ld(a.gin, (long long)q * g.nodes + pull_src<L>(nn[1], po, q), q, r);
            }
        }
    }
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        if (!use[e]) continue;
        tile[hy[e] * TX + hx[e]] = v[e];
        // the owner writes a bulk node; every halo holding a copied ring cell writes it
        if (store[e]) a.phi[nn[e]] = v[e];
    }
    __syncthreads();
    fused_node<L, ST, CT, SP, false, MODE_FULL, false, false, false, true>(a, 0u, tile, pf, pg);
}

// Inlet side, outside the disc: phi ghost <- boundary node (Solver.cu:621-633).
template <class CT>
__global__ void k_ghost_velocity(const CT* oldp, CT* newp, DevGrid g, int D, const unsigned char* flag, int side) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= plane_size(g, side)) return;
    const int a = side >> 1;
    const int t1 = (a == 0) ? 1 : 0;
    const int t2 = (a == 2) ? 1 : 2;
    int c[3];
    c[t1] = (int)(t % g.ext[t1]);
    c[t2] = (int)(t / g.ext[t1]);
    c[a] = (side & 1) ? g.dmax[a] + 1 : g.dmin[a] - 1;
    if (flag[t]) return;
    int b[3] = {c[0], c[1], c[2]};
    b[a] = (side & 1) ? g.dmax[a] : g.dmin[a];
    newp[g.idx(c[0], c[1], c[2])] = ghost_source(oldp, newp, g, D, b, 1);
}

// ------------------------------------------------------------ initialisation
// KernelInitializeDistributions with the interior stencils of
// deviceInitializer (NsAllen.cu:137-202, 897-908).  All nodes, both buffers;
// interior g encoded relative to phi, ghost cells absolute.
template <class L, class ST, class CT, class SP>
__global__ void k_init(StepArgs<ST, CT> a, ST* f2, ST* g2, CT L2, CT sigma) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const DevGrid& g = a.g;
    if (n >= g.nodes) return;
    int c[3];
    if (!g.coords(n, c)) return;
    const Phys<CT>& ph = a.ph;
    const CT c0 = a.phi[n];
    const CT rho = ph.rho_l + c0 * (ph.rho_h - ph.rho_l);
    CT p = CT(0);
    if (ph.case_type == CLIP_CASE_BUBBLE) p = CT(0) - c0 * sigma / L2 / (rho / CT(3));
    else if (ph.case_type == CLIP_CASE_DROP) p = CT(0) + c0 * sigma / L2 / (rho / CT(3));
    a.p[n] = p;
    for (int d = 0; d < L::D; ++d) a.uout[(long long)d * g.nodes + n] = CT(0);
    CT nr[3] = {0, 0, 0};
    const bool interior = g.interior(c, L::D);
    if (interior) {
        CT cc, dc[3], lap;
        phi_stencil<L, CT>(a.phi, g, n, cc, dc, lap);
        if constexpr (sizeof(CT) == 4) {
            CT rt, mag;
            normal_of<L>(dc, nr, rt, mag);
        } else {
            CT sq = dc[0] * dc[0] + dc[1] * dc[1];
            if (L::D == 3) sq += dc[2] * dc[2];
            const CT t = sqrt(sq) + CT(1e-32);
            for (int d = 0; d < L::D; ++d) nr[d] = dc[d] / t;
        }
    }
    const CT u0[3] = {0, 0, 0};
    const CT hw = (CT(1) - CT(4) * ((c0 - CT(0.5)) * (c0 - CT(0.5)))) / ph.W;
#pragma unroll
    for (int q = 0; q < L::Q; ++q) {
        const CT ga = gamma_u<L, CT>(q, u0);
        CT en = CT(L::ex(q)) * nr[0] + CT(L::ey(q)) * nr[1];
        if (L::D == 3) en += CT(L::ez(q)) * nr[2];
        const CT hlp = CT(L::w(q)) * (hw * en);
        const CT gv = c0 * (ga + CT(L::w(q))) - CT(0.5) * hlp;
        const CT fv = p * CT(L::w(q)) + ga;
        const long long i = (long long)q * g.nodes + n;
        SP::st(a.fout, i, q, fv, CT(0));
        SP::st(f2, i, q, fv, CT(0));
        const CT gr = (SP::uses_ref && interior) ? a.phi[ref_index<L>(g, a.bc, c, q)] : CT(0);
        SP::st(a.gout, i, q, gv, gr);
        SP::st(g2, i, q, gv, gr);
        if (a.bc.is_fc && interior) {
            store_prev<L, CT>(g, a.bc, a.prev_f_out, c, q, fv);
            store_prev<L, CT>(g, a.bc, a.prev_g_out, c, q, gv);
        }
    }
}

// Inlet populations inside the disc (JetBoundary, Solver.cu:578-619 and its
// y/z counterparts).  Every input is a ghost-cell value that no phase after
// initialisation rewrites, so the result is time-invariant and is computed
// once, into both buffers.
template <class L, class ST, class CT, class SP>
__global__ void k_inlet_init(DevGrid g, const unsigned char* flag, int side, const CT* phi, ST* fA, ST* gA, ST* fB,
                             ST* gB, const CT* Uplane) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= plane_size(g, side)) return;
    if (!flag[t]) return;
    const int ax = side >> 1;
    const int t1 = (ax == 0) ? 1 : 0;
    const int t2 = (ax == 2) ? 1 : 2;
    int c[3];
    c[t1] = (int)(t % g.ext[t1]);
    c[t2] = (int)(t / g.ext[t1]);
    c[ax] = (side & 1) ? g.dmax[ax] + 1 : g.dmin[ax] - 1;
    const long long n = g.idx(c[0], c[1], c[2]);
    const CT U[3] = {Uplane[3 * t + 0], Uplane[3 * t + 1], Uplane[3 * t + 2]};
    const CT cg = phi[n];
    const int inward = (side & 1) ? -1 : 1;
    CT feq[L::Q], geq[L::Q], Mt[3] = {0, 0, 0}, N = 0;
#pragma unroll
    for (int q = 1; q < L::Q; ++q) {
        const CT fa = gamma_u<L, CT>(q, U);
        feq[q] = fa;
        geq[q] = cg * (fa + CT(L::w(q)));
        if (L::e(q, ax) == 0) {
            const CT fq = SP::ld(fA, (long long)q * g.nodes + n, q, CT(0));
            const CT gq = SP::ld(gA, (long long)q * g.nodes + n, q, CT(0));
            for (int d = 0; d < L::D; ++d)
                if (d != ax) Mt[d] += CT(L::e(q, d)) * (fq - feq[q]);
            N += gq - geq[q];
        }
    }
    feq[0] = geq[0] = CT(0);
    const CT weight = CT(L::NIN - 1);
#pragma unroll
    for (int q = 1; q < L::Q; ++q) {
        if (L::e(q, ax) != inward) continue;
        const int b = L::opp(q);
        bool normal = true;
        CT corr = 0;
        for (int d = 0; d < L::D; ++d)
            if (d != ax) {
                corr += CT(L::e(q, d)) * Mt[d];
                if (L::e(q, d) != 0) normal = false;
            }
        const CT fb = SP::ld(fA, (long long)b * g.nodes + n, b, CT(0));
        const CT gb = SP::ld(gA, (long long)b * g.nodes + n, b, CT(0));
        const CT fv = fb + feq[q] - feq[b] - (CT(1) / weight) * corr;
        CT gv = geq[q] - (gb - geq[b]);
        if (!normal) gv -= N / weight;
        const long long i = (long long)q * g.nodes + n;
        SP::st(fA, i, q, fv, CT(0));
        SP::st(fB, i, q, fv, CT(0));
        SP::st(gA, i, q, gv, CT(0));
        SP::st(gB, i, q, gv, CT(0));
    }
}

// ------------------------------------------------------------ exports
// Post-streaming populations of every interior node (and p = sum f_post)
// into a dense CT array; ghost nodes copy the stored value.
template <class L, class ST, class CT, class SP, bool DIRECT>
__global__ void k_export_post(StepArgs<ST, CT> a, bool is_f, CT* out, CT* pout) {
    const ST* src = is_f ? a.fin : a.gin;
    const CT* const* prev = is_f ? a.prev_f_in : a.prev_g_in;
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const DevGrid& g = a.g;
    if (n >= g.nodes) return;
    int c[3];
    if (!g.coords(n, c)) return;
    const bool interior = g.interior(c, L::D);
    const CT* ref = is_f ? nullptr : a.phi_ref;
    CT sum = 0;
    for (int q = 0; q < L::Q; ++q) {
        CT v;
        if (!interior) v = SP::ld(src, (long long)q * g.nodes + n, q, CT(0));
        else if (DIRECT) v = SP::ld(src, (long long)q * g.nodes + n, q, gref<SP>(ref, ref_index<L>(g, a.bc, c, q)));
        else if (is_f) v = Puller<L, ST, CT, SP, false>{src, g, a.bc, a.ufc, prev, ref}.lvl3(c, q);
        else v = Puller<L, ST, CT, SP, true>{src, g, a.bc, a.ufc, prev, ref}.lvl3(c, q);
        out[(long long)q * g.nodes + n] = v;
        sum += v;
    }
    if (pout && interior) pout[n] = sum;
}

// Stored populations <-> dense CT array (ref = phi at interior nodes for g,
// nullptr for f).
template <class L, class ST, class CT, class SP>
__global__ void k_export_raw(const ST* src, CT* out, DevGrid g, DevBC bc, const CT* ref) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nodes * L::Q) return;
    const long long n = t % g.nodes;
    const int q = (int)(t / g.nodes);
    int c[3];
    if (!g.coords(n, c)) return;
    out[t] = SP::ld(src, t, q, g.interior(c, L::D) ? gref<SP>(ref, ref_index<L>(g, bc, c, q)) : CT(0));
}

template <class L, class ST, class CT, class SP>
__global__ void k_import_raw(ST* dst, const CT* in, DevGrid g, DevBC bc, const CT* ref) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nodes * L::Q) return;
    const long long n = t % g.nodes;
    const int q = (int)(t / g.nodes);
    int c[3];
    if (!g.coords(n, c)) return;
    SP::st(dst, t, q, in[t], g.interior(c, L::D) ? gref<SP>(ref, ref_index<L>(g, bc, c, q)) : CT(0));
}

// mu, grad phi and normal at interior nodes (ghosts left at 0).
template <class L, class CT>
__global__ void k_export_stencil(const CT* phi, DevGrid g, Phys<CT> ph, CT* mu, CT* dphi, CT* normal) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nodes) return;
    int c[3];
    if (!g.coords(n, c)) return;
    if (!g.interior(c, L::D)) {
        mu[n] = 0;
        for (int d = 0; d < L::D; ++d) dphi[d * g.nodes + n] = normal[d * g.nodes + n] = 0;
        return;
    }
    CT c0, dc[3], lap;
    phi_stencil<L, CT>(phi, g, n, c0, dc, lap);
    mu[n] = CT(4) * ph.beta * c0 * (c0 - CT(1)) * (c0 - CT(0.5)) - ph.kappa * lap;
    CT sq = dc[0] * dc[0] + dc[1] * dc[1];
    if (L::D == 3) sq += dc[2] * dc[2];
    const CT t = sqrt(sq) + CT(1e-32);
    for (int d = 0; d < L::D; ++d) {
        dphi[d * g.nodes + n] = dc[d];
        normal[d * g.nodes + n] = dc[d] / t;
    }
}

// ------------------------------------------------------------ reductions
// FP64 sum and non-finite count of a scalar field over interior nodes
// (Reporter::sum, Reporter.cu:50-114): wave64 shuffles, then LDS across the
// block's waves, one partial per block; a second launch folds the partials.
template <class CT>
__global__ __launch_bounds__(256) void k_reduce_interior(const CT* x, DevGrid g, double* part, unsigned* nonfinite) {
    __shared__ double ws[4];
    __shared__ unsigned wn[4];
    const long long nx = g.dmax[0] - g.dmin[0] + 1;
    const long long ny = g.dmax[1] - g.dmin[1] + 1;
    const long long nz = g.dmax[2] - g.dmin[2] + 1;
    const long long total = nx * ny * nz;
    double acc = 0.0;
    unsigned bad = 0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % nx) + g.dmin[0];
        const int j = (int)((t / nx) % ny) + g.dmin[1];
        const int k = (int)(t / (nx * ny)) + g.dmin[2];
        const double v = (double)x[g.idx(i, j, k)];
        if (!isfinite(v)) bad++;
        else acc += v;
    }
    for (int off = 32; off > 0; off >>= 1) {
        acc += __shfl_xor(acc, off, 64);
        bad += __shfl_xor(bad, off, 64);
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        ws[wave] = acc;
        wn[wave] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        unsigned b = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            s += ws[w];
            b += wn[w];
        }
        part[blockIdx.x] = s;
        if (b) atomicAdd(nonfinite, b);
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_reduce_final(const T* part, int nparts, T* out) {
    __shared__ T ws[4];
    T acc = 0.0;
    for (int t = threadIdx.x; t < nparts; t += blockDim.x) acc += part[t];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        T s = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += ws[w];
        out[0] = s;
    }
}

}  // namespace clipamd
