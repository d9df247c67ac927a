// solver_impl.hpp — SolverImpl<Lattice, StorageT, ComputeT, StorePolicy>:
// device buffers of one handle and the launch sequence of every entry point.
//
// State of a handle between calls (after clip_init_fresh or clip_step):
//   pop[cur]       f*, g* of the pending collision (the next step pulls it)
//   pop[1-cur]     f*, g* the last step pulled (the reference's dev_f / dev_g);
//                  right after initialisation or a resume it holds f_post directly
//   phib[ph]       phi of the last macroscopic step with the ghost ring its
//                  stencils saw; pop[cur] is encoded relative to it
//   phib[1-ph]     the phi before (pop[1-cur] is encoded relative to it)
//   u[ucur]        u of the last macroscopic step; u[1-ucur] the one before
//   prev[pcur]     outflow planes read by the next pull
// One step: ghost refresh (phib[ph] -> ghosts of phib[1-ph]), k_phase,
// k_fused (FULL), then every index flips.
#pragma once

#include "geometry.hpp"
#include "impl.hpp"
#include "adjoint.hpp"
#include "adjoint_maps.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

namespace clipamd {

template <class L, class ST, class CT, class SP>
struct SolverImpl : Impl {
    static constexpr int QL = L::Q;
    // Whole-row bulk addressing (block-uniform row base + 32-bit lane offset)
    // for one-row blocks.  Measured on the 256^3 bubble: FP64 k_fused 2.586 ->
    // 2.566 ms; FP32 1.307 -> 1.367 ms and the 16-bit storage unchanged, so
    // only the FP64 instantiation uses it.
    static constexpr bool kRowAddressing = sizeof(CT) == 8;
    // 16-bit storage: the bulk phase sum on x-adjacent node pairs
    // (k_phase_pair), one 32-bit access per pulled g of both nodes.  Measured
    // on the 256^3 bubble: 0.190 -> 0.164-0.184 ms.  The same pairing of k_fused
    // (two nodes per lane: 192 VGPRs; or the f pair and the g pair on
    // adjacent lanes with a lane swap) was slower than one node per lane
    // (1.39 and 1.05 ms against 0.83 ms), so k_fused is not paired.
    static constexpr bool kPairs = sizeof(ST) == 2;
    clip_config cfg;
    DevGrid g;
    DevBC bc;
    Phys<CT> ph;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    ST* fbuf[2] = {nullptr, nullptr};
    ST* gbuf[2] = {nullptr, nullptr};
    CT* phib[2] = {nullptr, nullptr};
    CT* u[2] = {nullptr, nullptr};
    CT* p = nullptr;
    CT* prevf[2][6] = {};
    CT* prevg[2][6] = {};
    unsigned char* vflag[6] = {};
    CT* scratch = nullptr;        // export staging (max(Q, D) * nodes)
    CT* shell_f = nullptr;        // boundary-layer f_post / g_post (Q * shell nodes)
    CT* shell_g = nullptr;
    double* red_part = nullptr;   // reduction partials
    unsigned* red_bad = nullptr;
    double* red_out = nullptr;
    int cur = 1, ucur = 0, pcur = 0, phc = 0;
    unsigned tbx = BX, tby = BY;  // block shape of the bulk launches (tbx * tby = BX * BY)
    bool timing_on = false;
    bool row_align = true;
    hipEvent_t ev[8] = {};
    double t_total = 0, t_ghost = 0, t_phase = 0, t_fused = 0;
    std::vector<void*> allocations;

    // Node array of `count` elements (a multiple of the padded node count),
    // offset so that element dmin = 1 of every row starts a 128-byte line.
    template <class T>
    T* alloc_grid(size_t count) {
        const size_t lead = row_align ? 128 / sizeof(T) - 1 : 0;
        return alloc<T>(count + lead + 1, lead);
    }

    template <class T>
    T* alloc(size_t count, size_t lead = 0) {
        void* ptr = nullptr;
        if (hipMalloc(&ptr, count * sizeof(T)) != hipSuccess) {
            (void)hipGetLastError();
            throw std::bad_alloc();
        }
        allocations.push_back(ptr);
        bytes += count * sizeof(T);
        return (T*)ptr + lead;
    }

    // Host arrays crossing the Impl interface are dense (x fastest over
    // ext[0], ghosts included); device node arrays use the padded stride.
    size_t hn() const { return (size_t)g.ext[0] * g.ext[1] * g.ext[2]; }
    size_t hidx(int i, int j, int k) const { return (size_t)i + (size_t)g.ext[0] * ((size_t)j + (size_t)g.ext[1] * k); }
    // Visits every node row: fn(dense offset of (0, j, k), device offset of (0, j, k)).
    template <class Fn>
    void for_rows(Fn fn) const {
        for (int k = 0; k < g.ext[2]; ++k)
            for (int j = 0; j < g.ext[1]; ++j) fn(hidx(0, j, k), (size_t)g.idx(0, j, k));
    }
    template <class V, class W>
    void dense_to_dev(const V& in, W& out, int ndof) const {
        out.assign((size_t)g.nodes * ndof, 0);
        for (int d = 0; d < ndof; ++d)
            for_rows([&](size_t h, size_t dv) {
                for (int i = 0; i < g.ext[0]; ++i) out[(size_t)d * g.nodes + dv + i] = in[(size_t)d * hn() + h + i];
            });
    }
    template <class V, class W>
    void dev_to_dense(const V& in, W& out, int ndof) const {
        out.assign(hn() * ndof, 0);
        for (int d = 0; d < ndof; ++d)
            for_rows([&](size_t h, size_t dv) {
                for (int i = 0; i < g.ext[0]; ++i) out[(size_t)d * hn() + h + i] = in[(size_t)d * g.nodes + dv + i];
            });
    }

    explicit SolverImpl(const clip_config& c) : cfg(c) {
        D = L::D;
        Q = L::Q;
        if (c.dim != L::D) throw ConfigError("lattice/dimension mismatch");
        // Row stride padded so that every row's first interior node starts a
        // 128-byte line in every node array (with the offsets of alloc_grid):
        // a row's collision output then covers whole lines, which no other
        // block (on another XCD's L2) shares.  Measured on a 19-plane 16-bit
        // pull kernel on the 256^3 grid: 3.76 -> 4.24 TB/s.
        // CLIP_AMD_ROW_ALIGN=0 keeps the dense stride (ext_x).
        {
            const char* env = std::getenv("CLIP_AMD_ROW_ALIGN");
            row_align = !(env && env[0] == '0');
            g = grid_of_config(c, L::D, row_align ? 128 / (long long)std::min(sizeof(ST), sizeof(CT)) : 1);
        }
        for (int a = 0; a < L::D; ++a)
            if (g.dmax[a] - g.dmin[a] + 1 < 3) throw ConfigError("the interior needs at least 3 nodes per axis");
        {
            // rows split into equal chunks of at most 256 nodes, each rounded
            // up to whole waves
            const int nx = g.dmax[0] - g.dmin[0] + 1;
            const int chunks = (nx + BX * BY - 1) / (BX * BY);
            const int per = (nx + chunks - 1) / chunks;
            tbx = (unsigned)((per + 63) / 64 * 64);
            tby = std::max(1u, (unsigned)(BX * BY) / tbx);
        }
        bc = bc_of_config(c, L::D);
        ph.rho_h = (CT)c.rho_h;
        ph.rho_l = (CT)c.rho_l;
        ph.tau_h = (CT)c.tau_h;
        ph.tau_l = (CT)c.tau_l;
        ph.beta = (CT)c.beta;
        ph.kappa = (CT)c.kappa;
        ph.mobility = (CT)c.mobility;
        ph.W = (CT)c.interface_width;
        for (int d = 0; d < 3; ++d) ph.grav[d] = (d < L::D) ? (CT)c.gravity[d] : CT(0);
        ph.case_type = c.case_type;

        HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        own_stream = true;
        for (auto& e : ev) HIP_CHECK(hipEventCreate(&e));
        const size_t nodes = (size_t)g.nodes;
        for (int b = 0; b < 2; ++b) {
            fbuf[b] = alloc_grid<ST>(nodes * QL);
            gbuf[b] = alloc_grid<ST>(nodes * QL);
            u[b] = alloc_grid<CT>(nodes * L::D);
            phib[b] = alloc_grid<CT>(nodes);
        }
        p = alloc_grid<CT>(nodes);
        scratch = alloc_grid<CT>(nodes * std::max(QL, L::D));
        if (shell_nodes() > 0) {
            if (shell_nodes() * QL >= (1LL << 31)) throw ConfigError("boundary layer too large");
            shell_f = alloc<CT>((size_t)shell_nodes() * QL);
            shell_g = alloc<CT>((size_t)shell_nodes() * QL);
        }
        red_part = alloc<double>(1024);
        red_bad = alloc<unsigned>(1);
        red_out = alloc<double>(1);
        for (int s = 0; s < 2 * L::D; ++s) {
            if (c.bc_type[s] == CLIP_BC_FREE_CONVECT)
                for (int b = 0; b < 2; ++b) {
                    prevf[b][s] = alloc<CT>(plane_size(g, s) * QL);
                    prevg[b][s] = alloc<CT>(plane_size(g, s) * QL);
                    HIP_CHECK(hipMemset(prevf[b][s], 0, plane_size(g, s) * QL * sizeof(CT)));
                    HIP_CHECK(hipMemset(prevg[b][s], 0, plane_size(g, s) * QL * sizeof(CT)));
                }
            if (c.bc_type[s] == CLIP_BC_VELOCITY) {
                const long long ps = plane_size(g, s);
                std::vector<unsigned char> fl(ps);
                const int a = s / 2;
                const int t1 = (a == 0) ? 1 : 0, t2 = (a == 2) ? 1 : 2;
                for (long long t = 0; t < ps; ++t) {
                    int cc[3];
                    cc[t1] = (int)(t % g.ext[t1]);
                    cc[t2] = (int)(t / g.ext[t1]);
                    cc[a] = (s & 1) ? g.dmax[a] + 1 : g.dmin[a] - 1;
                    const double z = (L::D == 3) ? (double)cc[2] : 0.0;
                    fl[t] = sdf(cfg, 0, cc[0], cc[1], z) <= 0 ? 1 : 0;
                }
                vflag[s] = alloc<unsigned char>(ps);
                HIP_CHECK(hipMemcpy(vflag[s], fl.data(), ps, hipMemcpyHostToDevice));
            }
            bc.vflag[s] = vflag[s];
        }
        for (int s = 2 * L::D; s < 6; ++s) bc.vflag[s] = nullptr;
        for (int b = 0; b < 2; ++b) {
            HIP_CHECK(hipMemsetAsync(phib[b], 0, nodes * sizeof(CT), stream));
            HIP_CHECK(hipMemsetAsync(u[b], 0, nodes * L::D * sizeof(CT), stream));
        }
        HIP_CHECK(hipMemsetAsync(p, 0, nodes * sizeof(CT), stream));
        HIP_CHECK(hipStreamSynchronize(stream));
    }

    ~SolverImpl() override {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& e : graphs) (void)hipGraphExecDestroy(e.second);
        adj_free();
        for (void* ptr : allocations) (void)hipFree(ptr);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (own_stream && stream) (void)hipStreamDestroy(stream);
    }

    void set_stream(hipStream_t s) override {
        HIP_CHECK(hipStreamSynchronize(stream));
        if (s == nullptr) {
            if (!own_stream) {
                HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
                own_stream = true;
            }
            return;
        }
        if (own_stream) HIP_CHECK(hipStreamDestroy(stream));
        stream = s;
        own_stream = false;
    }

    void sync() override { HIP_CHECK(hipStreamSynchronize(stream)); }
    void* stream_handle() override { return (void*)stream; }

    // Arguments of a pass that pulls pop[pull_buf] (encoded relative to
    // phi_ref) into pop[1-pull_buf], u[uin] -> u[1-uin], outflow planes
    // prev[prev_in] -> prev[1-prev_in], stencils from `phi_new`.
    StepArgs<ST, CT> args(int pull_buf, int uin, int prev_in, CT* phi_new, const CT* phi_ref) {
        StepArgs<ST, CT> a;
        a.g = g;
        a.bc = bc;
        a.shell = make_shell<L>(g, bc);
        a.ph = ph;
        a.fin = fbuf[pull_buf];
        a.gin = gbuf[pull_buf];
        a.fout = fbuf[1 - pull_buf];
        a.gout = gbuf[1 - pull_buf];
        a.phi = phi_new;
        a.phi_ref = phi_ref;
        a.uin = u[uin];
        a.ufc = u[uin];
        a.uout = u[1 - uin];
        a.ucol = u[uin];
        a.p = p;
        a.shell_f = shell_f;
        a.shell_g = shell_g;
        for (int s = 0; s < 6; ++s) {
            a.prev_f_in[s] = prevf[prev_in][s];
            a.prev_f_out[s] = prevf[1 - prev_in][s];
            a.prev_g_in[s] = prevg[prev_in][s];
            a.prev_g_out[s] = prevg[1 - prev_in][s];
        }
        return a;
    }

    // Grid over the interior box shrunk by `inset` along each axis that is
    // not periodic on both sides (the bulk region of the pulling kernels).
    dim3 grid_box(int inset) const {
        int n[3];
        for (int a = 0; a < 3; ++a) {
            const int in = (a < L::D && inset && !bc.both_periodic[a]) ? 1 : 0;
            n[a] = g.dmax[a] - g.dmin[a] + 1 - 2 * in;
        }
        if (n[0] <= 0 || n[1] <= 0 || n[2] <= 0) return dim3(0, 0, 0);
        return dim3((n[0] + tbx - 1) / tbx, (n[1] + tby - 1) / tby, n[2]);
    }
    dim3 grid_interior() const { return grid_box(0); }
    // The bulk box of grid_box(1) as rows of x-adjacent node pairs.
    PairGrid pair_grid() const {
        PairGrid pg{};
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            const int in = (a < L::D && !bc.both_periodic[a]) ? 1 : 0;
            lo[a] = g.dmin[a] + in;
            hi[a] = g.dmax[a] - in;
        }
        if (hi[0] < lo[0] || hi[1] < lo[1] || hi[2] < lo[2]) return pg;
        pg.xlo = lo[0];
        pg.xhi = hi[0];
        pg.ylo = lo[1];
        pg.ny = hi[1] - lo[1] + 1;
        pg.zlo = lo[2];
        pg.per_row = (unsigned)((hi[0] - lo[0] + 2) / 2);
        const unsigned long long total = (unsigned long long)pg.per_row * pg.ny * (hi[2] - lo[2] + 1);
        if (total > 0x7fffffffull) throw ConfigError("bulk too large for the pair launch (2^31 pairs)");
        pg.total = (unsigned)total;
        return pg;
    }
    unsigned pair_blocks(const PairGrid& pg) const { return (pg.total + BX * BY - 1) / (BX * BY); }
    dim3 block_shape() const { return dim3(tbx, tby); }
    long long shell_nodes() const {
        const ShellDesc sd = make_shell<L>(g, bc);
        long long t = 0;
        for (int k = 0; k < sd.na; ++k) t += sd.count[k];
        return t;
    }
    unsigned shell_blocks() const { return (unsigned)((shell_nodes() + BX * BY - 1) / (BX * BY)); }

    void launch_check(const char* what) {
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) throw HipError(std::string("launch of ") + what + ": " + hipGetErrorString(e));
    }

    // Bulk and boundary-layer launches of the pulling step kernels.
    template <int MODE>
    void launch_fused_pull(const StepArgs<ST, CT>& a) {
        const dim3 gb = grid_box(1);
        const long long S = shell_nodes();
        if constexpr (kCarryShell) {
            if (gb.x && S) {
                // the boundary layer's nodes ride on the bulk launch as extra z-slices
                dim3 grid = gb;
                const unsigned long long tpb = (unsigned long long)tbx * tby,
                                         per_slice = (unsigned long long)gb.x * gb.y;
                const unsigned long long blocks = ((unsigned long long)S + tpb - 1) / tpb;
                grid.z += (unsigned)((blocks + per_slice - 1) / per_slice);
                bool rows = false;
                if constexpr (kRowAddressing) {
                    rows = tby == 1;
                    if (rows)
                        k_fused<L, ST, CT, SP, false, MODE, false, true, true><<<grid, block_shape(), 0, stream>>>(a, gb.z);
                }
                if (!rows)
                    k_fused<L, ST, CT, SP, false, MODE, false, false, true><<<grid, block_shape(), 0, stream>>>(a, gb.z);
                launch_check("k_fused");
                return;
            }
        }
        bool done = false;
        if constexpr (kRowAddressing) {
            done = gb.x && tby == 1;
            if (done) k_fused<L, ST, CT, SP, false, MODE, false, true><<<gb, block_shape(), 0, stream>>>(a, 0u);
        }
        if (gb.x && !done) k_fused<L, ST, CT, SP, false, MODE, false><<<gb, block_shape(), 0, stream>>>(a, 0u);
        if (S) k_fused<L, ST, CT, SP, false, MODE, true><<<shell_blocks(), BX * BY, 0, stream>>>(a, 0u);
        launch_check("k_fused");
    }
    // phi^{n+1} of the bulk (k_phase) and of the boundary layer (k_shell_pull,
    // which also leaves the layer's f_post / g_post for k_fused).
    // The bulk phase launch.  A ghost job (gj.total > 0, from ghost_job) rides
    // on it as extra blocks, or runs as its own k_ghost_resolve launch when
    // there is no bulk launch.
    void launch_phase_pull(const StepArgs<ST, CT>& a, GhostJob<CT> gj = GhostJob<CT>{}) {
        const dim3 gb = grid_box(1);
        bool done = false, ghost_done = gj.total == 0;
        // the box grid with enough extra z-slices of blocks for the ghost cells
        auto with_ghosts = [&](dim3 grid) {
            if (!gj.total) return grid;
            const unsigned long long tpb = (unsigned long long)tbx * tby, per_slice = (unsigned long long)grid.x * grid.y;
            const unsigned long long blocks = ((unsigned long long)gj.total + tpb - 1) / tpb;
            gj.first = grid.z;
            grid.z += (unsigned)((blocks + per_slice - 1) / per_slice);
            ghost_done = true;
            return grid;
        };
        if constexpr (kRowAddressing) {
            done = gb.x && tby == 1;
            if (done) {
                const dim3 grid = with_ghosts(gb);
                k_phase<L, ST, CT, SP, false, false, true><<<grid, block_shape(), 0, stream>>>(a, gj);
            }
        }
        if constexpr (kPairs) {
            const PairGrid pg = pair_grid();
            if (pg.total) {
                unsigned blocks = pair_blocks(pg);
                if (gj.total) {
                    gj.first = blocks;
                    blocks += (unsigned)((gj.total + BX * BY - 1) / (BX * BY));
                    ghost_done = true;
                }
                k_phase_pair<L, ST, SP><<<blocks, BX * BY, 0, stream>>>(a, pg, gj);
            }
            done = true;
        }
        if (gb.x && !done) {
            const dim3 grid = with_ghosts(gb);
            k_phase<L, ST, CT, SP, false, false><<<grid, block_shape(), 0, stream>>>(a, gj);
        }
        launch_check("k_phase");
        if (!ghost_done) {
            k_ghost_resolve<CT><<<(unsigned)((gj.total + 255) / 256), 256, 0, stream>>>(gj.oldp, gj.newp, g, bc, L::D,
                                                                                       gj.total, 1, 1, 0);
            launch_check("k_ghost_resolve");
        }
        if (shell_blocks()) {
            launch_shell_pull(a);
            k_shell_phi<L, ST, CT><<<shell_blocks(), BX * BY, 0, stream>>>(a);
            launch_check("k_shell_phi");
        }
    }
    void launch_shell_pull(const StepArgs<ST, CT>& a) {
        k_shell_pull<L, ST, CT, SP><<<dim3(shell_blocks(), QL), BX * BY, 0, stream>>>(a);
        launch_check("k_shell_pull");
    }

    // 2D steps with plain storage, x periodic on both sides and no inlet or
    // free-convect side run the phase sum, the ghost ring and the collision
    // as one k_step_tile launch: 1 = every side periodic; 2 = y bounded, its
    // boundary layer (the two rows next to the y sides) pulled first by
    // k_shell_pull / k_shell_phi; 0 = the separate launches.
    int tile_mode() const {
        if constexpr (L::D != 2 || SP::uses_ref) {
            return 0;
        } else {
            if (bc.is_velocity || bc.is_fc || !bc.both_periodic[0]) return 0;
            if (bc.both_periodic[1]) return 1;
            const ShellDesc sd = make_shell<L>(g, bc);
            return (sd.na == 1 && sd.axes[0] == 1) ? 2 : 0;
        }
    }
    bool tile_step() const { return tile_mode() == 1; }
    void launch_step_tile(const StepArgs<ST, CT>& a, const CT* phi_old, int mode) {
        if constexpr (L::D == 2 && !SP::uses_ref) {
            const dim3 grid((unsigned)((g.dmax[0] - g.dmin[0] + BX) / BX),
                            (unsigned)((g.dmax[1] - g.dmin[1] + BY) / BY), 1);
            if (mode == 2) {
                launch_shell_pull(a);
                k_shell_phi<L, ST, CT><<<shell_blocks(), BX * BY, 0, stream>>>(a);
                launch_check("k_shell_phi");
                k_step_tile<L, ST, CT, SP, true><<<grid, dim3(BX, BY), 0, stream>>>(a, phi_old);
            } else {
                k_step_tile<L, ST, CT, SP, false><<<grid, dim3(BX, BY), 0, stream>>>(a, phi_old);
            }
            launch_check("k_step_tile");
        }
    }

    // Without inlet sides the ghost ring is one k_ghost_resolve pass, which
    // can ride on the phase launch (launch_phase_pull); total = 0 otherwise.
    GhostJob<CT> ghost_job(const CT* phi_old, CT* phi_new) const {
        if (bc.is_velocity) return GhostJob<CT>{};
        long long total = 0;
        for (int s = 0; s < 2 * L::D; ++s) total += plane_size(g, s);
        return GhostJob<CT>{phi_old, phi_new, total, 0u};
    }
    // Ghost ring of phi_new from phi_old (Solver.cu:75-141, 550-921, 927-978).
    void ghost_refresh(const CT* phi_old, CT* phi_new) {
        long long total = 0;
        for (int s = 0; s < 2 * L::D; ++s) total += plane_size(g, s);
        const int tb = 256;
        const unsigned blocks = (unsigned)((total + tb - 1) / tb);
        if (!bc.is_velocity) {
            k_ghost_resolve<CT><<<blocks, tb, 0, stream>>>(phi_old, phi_new, g, bc, L::D, total, 1, 1, 0);
            launch_check("k_ghost_resolve");
            return;
        }
        k_ghost_resolve<CT><<<blocks, tb, 0, stream>>>(phi_old, phi_new, g, bc, L::D, total, 1, 0, 0);
        launch_check("k_ghost_resolve");
        for (int s = 0; s < 2 * L::D; ++s) {
            if (cfg.bc_type[s] != CLIP_BC_VELOCITY) continue;
            const long long ps = plane_size(g, s);
            k_ghost_velocity<CT><<<(unsigned)((ps + tb - 1) / tb), tb, 0, stream>>>(phi_old, phi_new, g, L::D,
                                                                                 vflag[s], s);
            launch_check("k_ghost_velocity");
        }
        if (bc.mirror_gate) {
            k_ghost_resolve<CT><<<blocks, tb, 0, stream>>>(phi_old, phi_new, g, bc, L::D, total, 0, 1, 1);
            launch_check("k_ghost_resolve");
        }
    }

    // Incoming populations of the ghost cells inside each inlet disc
    // (JetBoundary, Solver.cu:578-619) from the tangential and outgoing ghost
    // populations of pop 0 and phi at the ghost cell, into both buffers.  Its
    // inputs are ghost values no phase of a step rewrites, so it runs at
    // initialisation and after an upload of populations or phi.
    void inlet_fill(const CT* phi) {
        const int tb = 256;
        for (int s = 0; s < 2 * L::D; ++s) {
            if (cfg.bc_type[s] != CLIP_BC_VELOCITY) continue;
            const long long ps = plane_size(g, s);
            std::vector<CT> U(3 * ps);
            const int ax = s / 2;
            const int t1 = (ax == 0) ? 1 : 0, t2 = (ax == 2) ? 1 : 2;
            for (long long t = 0; t < ps; ++t) {
                int cc[3];
                cc[t1] = (int)(t % g.ext[t1]);
                cc[t2] = (int)(t / g.ext[t1]);
                cc[ax] = (s & 1) ? g.dmax[ax] + 1 : g.dmin[ax] - 1;
                double Uv[3];
                inlet_velocity(cfg, s, cc, Uv);
                for (int d = 0; d < 3; ++d) U[3 * t + d] = (CT)Uv[d];
            }
            CT* dU = scratch;
            HIP_CHECK(hipMemcpyAsync(dU, U.data(), U.size() * sizeof(CT), hipMemcpyHostToDevice, stream));
            k_inlet_init<L, ST, CT, SP><<<(unsigned)((ps + tb - 1) / tb), tb, 0, stream>>>(
                g, vflag[s], s, phi, fbuf[0], gbuf[0], fbuf[1], gbuf[1], dU);
            launch_check("k_inlet_init");
            HIP_CHECK(hipStreamSynchronize(stream));
        }
    }

    void init_fresh() override {
        // Host phi over every node including ghosts (NsAllen.cu:768-824).
        std::vector<CT> h(g.nodes);
        for (int k = 0; k < g.ext[2]; ++k)
            for (int j = 0; j < g.ext[1]; ++j)
                for (int i = 0; i < g.ext[0]; ++i) h[g.idx(i, j, k)] = (CT)initial_phi(cfg, i, j, k);
        init_with_phi(h);
    }

    void init_phi(const std::vector<double>& phi) override {
        if (phi.size() != hn()) throw ConfigError("phi must hold every node, ghosts included");
        std::vector<CT> h(g.nodes, CT(0));
        for (int k = 0; k < g.ext[2]; ++k)
            for (int j = 0; j < g.ext[1]; ++j)
                for (int i = 0; i < g.ext[0]; ++i) h[g.idx(i, j, k)] = (CT)phi[hidx(i, j, k)];
        init_with_phi(h);
    }

    // Initial state (KernelInitializeDistributions) from a phi given on every
    // node, then the first collision.
    void init_with_phi(const std::vector<CT>& h) {
        for (int b = 0; b < 2; ++b)
            HIP_CHECK(hipMemcpyAsync(phib[b], h.data(), h.size() * sizeof(CT), hipMemcpyHostToDevice, stream));
        init_from_phib();
    }
    // The same from phi already in both phi buffers.
    void init_from_phib() {
        // Distributions, p, u on all nodes of both buffers; outflow planes.
        StepArgs<ST, CT> a = args(1, 1, 1, phib[0], phib[0]);  // fout = pop 0, uout = u[0], prev_out = prev[0]
        const int tb = 256;
        k_init<L, ST, CT, SP><<<(unsigned)((g.nodes + tb - 1) / tb), tb, 0, stream>>>(
            a, fbuf[1], gbuf[1], (CT)(cfg.reference_length / 2.0), (CT)cfg.sigma);
        launch_check("k_init");
        inlet_fill(phib[0]);
        // First collision straight from the initial f_post = f (pop 0).
        StepArgs<ST, CT> c = args(0, 0, 1, phib[0], phib[0]);
        k_fused<L, ST, CT, SP, true, MODE_COLLIDE, false><<<grid_interior(), block_shape(), 0, stream>>>(c, 0u);
        launch_check("k_fused");
        HIP_CHECK(hipStreamSynchronize(stream));
        cur = 1;
        ucur = 0;
        pcur = 0;
        phc = 0;
        direct_old = true;
        old_ref_zero = false;
        direct_f_interior.clear();
        direct_g_interior.clear();
        step_count = 0;
        initialized = true;
    }

    // One step's launches from the current buffer indices, then every index
    // flips (the collision just written is pulled next).
    void one_step() {
        CT* phi_old = phib[phc];
        CT* phi_new = phib[1 - phc];
        StepArgs<ST, CT> a = args(cur, ucur, pcur, phi_new, phi_old);
        const bool tm = timing_on && !capturing;
        if (tm) HIP_CHECK(hipEventRecord(ev[0], stream));
        const int tiled = tile_mode();
        // the ghost ring: with the phase launch when it is one pass
        GhostJob<CT> gj{};
        if (!tiled) {
            gj = ghost_job(phi_old, phi_new);
            if (!gj.total) ghost_refresh(phi_old, phi_new);
        }
        if (tm) HIP_CHECK(hipEventRecord(ev[1], stream));
        if (!tiled) launch_phase_pull(a, gj);
        if (tm) HIP_CHECK(hipEventRecord(ev[2], stream));
        if (tiled) launch_step_tile(a, phi_old, tiled);
        else launch_fused_pull<MODE_FULL>(a);
        if (tm) {
            HIP_CHECK(hipEventRecord(ev[3], stream));
            HIP_CHECK(hipEventSynchronize(ev[3]));
            float m1, m2, m3;
            HIP_CHECK(hipEventElapsedTime(&m1, ev[0], ev[1]));
            HIP_CHECK(hipEventElapsedTime(&m2, ev[1], ev[2]));
            HIP_CHECK(hipEventElapsedTime(&m3, ev[2], ev[3]));
            t_ghost += m1;
            t_phase += m2;
            t_fused += m3;
            t_total += m1 + m2 + m3;
        }
        cur = 1 - cur;
        ucur = 1 - ucur;
        pcur = 1 - pcur;
        phc = 1 - phc;
        direct_old = false;
        old_ref_zero = false;
        step_count++;
    }

    // HIP graph of two steps: every index flips twice, so the graph of one
    // index state replays as many times as needed.  Its nodes are kernel
    // launches only (no memset / copy nodes).
    bool graph_on = false, capturing = false;
    std::vector<std::pair<int, hipGraphExec_t>> graphs;
    int index_key() const { return cur | (ucur << 1) | (pcur << 2) | (phc << 3); }
    hipGraphExec_t pair_graph() {
        const int key = index_key();
        for (auto& e : graphs)
            if (e.first == key) return e.second;
        const int c0 = cur, u0 = ucur, p0 = pcur, f0 = phc;
        const bool d0 = direct_old, z0 = old_ref_zero;
        const long long s0 = step_count;
        hipGraph_t gph = nullptr;
        capturing = true;
        HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        one_step();
        one_step();
        const hipError_t ec = hipStreamEndCapture(stream, &gph);
        capturing = false;
        HIP_CHECK(ec);
        hipGraphExec_t exec = nullptr;
        HIP_CHECK(hipGraphInstantiate(&exec, gph, nullptr, nullptr, 0));
        HIP_CHECK(hipGraphDestroy(gph));
        // capture records the launches without running them
        cur = c0;
        ucur = u0;
        pcur = p0;
        phc = f0;
        direct_old = d0;
        old_ref_zero = z0;
        step_count = s0;
        graphs.push_back({key, exec});
        return exec;
    }
    void enable_graph(bool on) override { graph_on = on; }

    void step(int n) override {
        t_total = t_ghost = t_phase = t_fused = 0;
        if (graph_on && !timing_on) {
            while (n >= 2) {
                HIP_CHECK(hipGraphLaunch(pair_graph(), stream));
                step_count += 2;
                direct_old = false;
                old_ref_zero = false;
                n -= 2;
            }
        }
        for (int s = 0; s < n; ++s) one_step();
    }

    void timing(double* total, double* gh, double* pha, double* fu) const override {
        if (total) *total = t_total;
        if (gh) *gh = t_ghost;
        if (pha) *pha = t_phase;
        if (fu) *fu = t_fused;
    }
    void enable_timing(bool on) override { timing_on = on; }

    template <class T>
    void download(const T* dev, size_t count, std::vector<double>& out) {
        std::vector<T> h(count);
        HIP_CHECK(hipMemcpyAsync(h.data(), dev, count * sizeof(T), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        out.resize(count);
        for (size_t i = 0; i < count; ++i) out[i] = (double)h[i];
    }

    // Node field of `ndof` components: device layout <-> dense host array.
    template <class T>
    void download_field(const T* dev, int ndof, std::vector<double>& out) {
        std::vector<double> raw;
        download(dev, (size_t)g.nodes * ndof, raw);
        dev_to_dense(raw, out, ndof);
    }
    template <class T>
    void upload_field(T* dev, const std::vector<double>& in, int ndof) {
        std::vector<double> d;
        dense_to_dev(in, d, ndof);
        upload(dev, d);
    }
    // Stored bytes of a population buffer: device layout <-> dense.
    void raw_to_dense(const std::vector<unsigned char>& in, std::vector<unsigned char>& out) const {
        out.assign(hn() * QL * sizeof(ST), 0);
        for (int q = 0; q < QL; ++q)
            for_rows([&](size_t h, size_t dv) {
                std::memcpy(&out[((size_t)q * hn() + h) * sizeof(ST)], &in[((size_t)q * g.nodes + dv) * sizeof(ST)],
                            (size_t)g.ext[0] * sizeof(ST));
            });
    }
    void raw_to_dev(const std::vector<unsigned char>& in, std::vector<unsigned char>& out) const {
        out.assign((size_t)g.nodes * QL * sizeof(ST), 0);
        for (int q = 0; q < QL; ++q)
            for_rows([&](size_t h, size_t dv) {
                std::memcpy(&out[((size_t)q * g.nodes + dv) * sizeof(ST)], &in[((size_t)q * hn() + h) * sizeof(ST)],
                            (size_t)g.ext[0] * sizeof(ST));
            });
    }

    // phi the pulled buffer pop[1-cur] is encoded with (nullptr = 0).
    const CT* old_ref() const { return old_ref_zero ? nullptr : phib[1 - phc]; }

    // Exports f_post/g_post (pull of the buffer the last step read, with the
    // u and outflow planes that pull used) into `scratch`; with is_f also
    // refreshes p = sum_q f_post.
    void export_post(bool is_f) {
        const int old = 1 - cur;
        StepArgs<ST, CT> a = args(old, 1 - ucur, 1 - pcur, phib[phc], old_ref());
        const int tb = 256;
        const unsigned blocks = (unsigned)((g.nodes + tb - 1) / tb);
        if (direct_old)  // p is the stored one (initial or resumed)
            k_export_post<L, ST, CT, SP, true><<<blocks, tb, 0, stream>>>(a, is_f, scratch, nullptr);
        else
            k_export_post<L, ST, CT, SP, false><<<blocks, tb, 0, stream>>>(a, is_f, scratch, is_f ? p : nullptr);
        launch_check("k_export_post");
    }

    void get_field(int field, std::vector<double>& out) override {
        const size_t nodes = (size_t)g.nodes;
        const int tb = 256;
        CT* phi = phib[phc];
        switch (field) {
            case CLIP_FIELD_PHI: download_field(phi, 1, out); return;
            case CLIP_FIELD_RHO: {
                download_field(phi, 1, out);
                for (auto& v : out) v = (double)(ph.rho_l + (CT)v * (ph.rho_h - ph.rho_l));
                return;
            }
            case CLIP_FIELD_U: download_field(u[ucur], L::D, out); return;
            case CLIP_FIELD_P: {
                if (!direct_old) export_post(true);
                download_field(p, 1, out);
                return;
            }
            case CLIP_FIELD_MU:
            case CLIP_FIELD_DPHI:
            case CLIP_FIELD_NORMAL: {
                CT* mu = scratch;
                CT* dp = scratch + nodes;
                CT* nr = scratch + nodes * (1 + L::D);
                k_export_stencil<L, CT><<<(unsigned)((nodes + tb - 1) / tb), tb, 0, stream>>>(phi, g, ph, mu, dp, nr);
                launch_check("k_export_stencil");
                if (field == CLIP_FIELD_MU) download_field(mu, 1, out);
                else if (field == CLIP_FIELD_DPHI) download_field(dp, L::D, out);
                else download_field(nr, L::D, out);
                return;
            }
            case CLIP_FIELD_F:
            case CLIP_FIELD_G: {
                const int old = 1 - cur;
                const bool is_f = field == CLIP_FIELD_F;
                const ST* src = is_f ? fbuf[old] : gbuf[old];
                const long long count = (long long)nodes * QL;
                k_export_raw<L, ST, CT, SP><<<(unsigned)((count + tb - 1) / tb), tb, 0, stream>>>(
                    src, scratch, g, bc, is_f ? nullptr : old_ref());
                launch_check("k_export_raw");
                download_field(scratch, QL, out);
                const auto& keep = is_f ? direct_f_interior : direct_g_interior;
                if (direct_old && !keep.empty()) out = keep;
                if (!direct_old) fill_periodic_ghosts(out, QL);
                return;
            }
            case CLIP_FIELD_F_POST:
            case CLIP_FIELD_G_POST: {
                export_post(field == CLIP_FIELD_F_POST);
                download_field(scratch, QL, out);
                return;
            }
        }
        throw ConfigError("unknown field");
    }

    // Periodic ghost cells are never read by the pull (it wraps), so they are
    // filled here the way kernelPeriodicBoundary would (Solver.cu:75-141).
    void fill_periodic_ghosts(std::vector<double>& a, int ndof) {
        for (int ax = 0; ax < L::D; ++ax) {
            if (!bc.periodic[ax]) continue;
            for (int k = 0; k < g.ext[2]; ++k)
                for (int j = 0; j < g.ext[1]; ++j)
                    for (int i = 0; i < g.ext[0]; ++i) {
                        int c[3] = {i, j, k};
                        if (c[ax] != 0 && c[ax] != g.ext[ax] - 1) continue;
                        int s[3] = {i, j, k};
                        s[ax] = (c[ax] == 0) ? g.dmax[ax] : g.dmin[ax];
                        for (int q = 0; q < ndof; ++q)
                            a[(size_t)q * hn() + hidx(i, j, k)] = a[(size_t)q * hn() + hidx(s[0], s[1], s[2])];
                    }
        }
    }

    template <class T>
    void upload(T* dev, const std::vector<double>& in) {
        std::vector<T> h(in.size());
        for (size_t i = 0; i < in.size(); ++i) h[i] = (T)in[i];
        HIP_CHECK(hipMemcpyAsync(dev, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
    }

    // Encodes populations (dense, SoA) into a population buffer; `ref` = phi
    // at interior nodes for g (nullptr: absolute).
    void upload_pop(ST* dev, const std::vector<double>& in, const CT* ref) {
        const long long count = (long long)g.nodes * QL;
        upload_field(scratch, in, QL);
        const int tb = 256;
        k_import_raw<L, ST, CT, SP><<<(unsigned)((count + tb - 1) / tb), tb, 0, stream>>>(dev, scratch, g, bc, ref);
        launch_check("k_import_raw");
        HIP_CHECK(hipStreamSynchronize(stream));
    }

    // Replaces the interior of a population array (SoA, ghosts included) by
    // that of `src`.
    void replace_interior(std::vector<double>& dst, const std::vector<double>& src) const {
        for (int k = g.dmin[2]; k <= g.dmax[2]; ++k)
            for (int j = g.dmin[1]; j <= g.dmax[1]; ++j)
                for (int i = g.dmin[0]; i <= g.dmax[0]; ++i)
                    for (int q = 0; q < QL; ++q) {
                        const size_t t = (size_t)q * hn() + hidx(i, j, k);
                        dst[t] = src[t];
                    }
    }

    // Direct state: the handle holds the reference's device arrays as they
    // stand between two solve() calls.  pop[0] = dev_f with its interior
    // replaced by f_post (stored absolutely, read in place), pop[1] = dev_f
    // (its ghost cells are what later pulls read; the collision rewrites the
    // interior); the same for g.  p, u[ucur] and phib[phc] are dev_p, dev_vel
    // and dev_c; the pending collision is recomputed from them.
    void set_direct(const std::vector<double>& fpost, const std::vector<double>& gpost, const std::vector<double>& df,
                    const std::vector<double>& dg) {
        std::vector<double> a0 = df, b0 = dg;
        replace_interior(a0, fpost);
        replace_interior(b0, gpost);
        upload_pop(fbuf[0], a0, nullptr);
        upload_pop(gbuf[0], b0, nullptr);
        upload_pop(fbuf[1], df, nullptr);
        upload_pop(gbuf[1], dg, nullptr);
        cur = 1;
        direct_old = true;
        old_ref_zero = true;
        direct_f_interior = df;
        direct_g_interior = dg;
    }

    // Moves the handle into the direct state with absolutely stored f_post /
    // g_post (exact for FP64 / FP32 storage; 16-bit storage re-encodes the
    // pending f_post / g_post once without the phi reference).
    void materialise() {
        if (direct_old && old_ref_zero) return;
        std::vector<double> fpost, gpost, df, dg;
        get_field(CLIP_FIELD_F_POST, fpost);  // first: refreshes p = sum f_post
        get_field(CLIP_FIELD_G_POST, gpost);
        get_field(CLIP_FIELD_F, df);
        get_field(CLIP_FIELD_G, dg);
        set_direct(fpost, gpost, df, dg);
    }

    // Writes one of the reference's device arrays.  The arrays keep the
    // reference's roles: dev_f / dev_g are overwritten by the next collision
    // except for their ghost cells, which later pulls read; f_post / g_post,
    // p, u and phi feed the next collision (whose mu, grad phi and normal
    // are recomputed from the phi given).  Nothing else is recomputed:
    // clip_macroscopic does that, as NSAllen::macroscopic does.
    void set_field(int field, const std::vector<double>& in) override {
        if (field != CLIP_FIELD_PHI && field != CLIP_FIELD_U && field != CLIP_FIELD_P && field != CLIP_FIELD_F &&
            field != CLIP_FIELD_G && field != CLIP_FIELD_F_POST && field != CLIP_FIELD_G_POST)
            throw ConfigError("field cannot be set (derived from phi)");
        materialise();
        switch (field) {
            case CLIP_FIELD_PHI:
                // Both phi buffers: ghost cells that no step rewrites (inside an
                // inlet disc, sides without a mirror) are read from whichever
                // buffer is current.  The direct state stores f_post without a
                // phi reference, so the other buffer is free.
                upload_field(phib[phc], in, 1);
                HIP_CHECK(hipMemcpyAsync(phib[1 - phc], phib[phc], (size_t)g.nodes * sizeof(CT),
                                         hipMemcpyDeviceToDevice, stream));
                break;
            case CLIP_FIELD_U: upload_field(u[ucur], in, L::D); break;
            case CLIP_FIELD_P: upload_field(p, in, 1); break;
            case CLIP_FIELD_F_POST:
            case CLIP_FIELD_G_POST: {
                const bool is_f = field == CLIP_FIELD_F_POST;
                std::vector<double> a0 = is_f ? direct_f_interior : direct_g_interior;
                replace_interior(a0, in);
                upload_pop(is_f ? fbuf[0] : gbuf[0], a0, nullptr);
                break;
            }
            case CLIP_FIELD_F:
            case CLIP_FIELD_G: {
                const bool is_f = field == CLIP_FIELD_F;
                std::vector<double> post;
                get_field(is_f ? CLIP_FIELD_F_POST : CLIP_FIELD_G_POST, post);
                std::vector<double> a0 = in;
                replace_interior(a0, post);
                upload_pop(is_f ? fbuf[0] : gbuf[0], a0, nullptr);
                upload_pop(is_f ? fbuf[1] : gbuf[1], in, nullptr);
                (is_f ? direct_f_interior : direct_g_interior) = in;
                break;
            }
        }
        if (bc.is_velocity && (field == CLIP_FIELD_F || field == CLIP_FIELD_G || field == CLIP_FIELD_PHI)) {
            inlet_fill(phib[phc]);
            // keep the host copies of dev_f / dev_g in step with the ghosts
            std::vector<double> raw;
            for (int is_f = 0; is_f < 2; ++is_f) {
                const int tb = 256;
                const long long count = (long long)g.nodes * QL;
                k_export_raw<L, ST, CT, SP><<<(unsigned)((count + tb - 1) / tb), tb, 0, stream>>>(
                    is_f ? fbuf[1] : gbuf[1], scratch, g, bc, nullptr);
                launch_check("k_export_raw");
                download_field(scratch, QL, raw);
                std::vector<double>& keep = is_f ? direct_f_interior : direct_g_interior;
                for (int k = 0; k < g.ext[2]; ++k)
                    for (int j = 0; j < g.ext[1]; ++j)
                        for (int i = 0; i < g.ext[0]; ++i) {
                            const int c[3] = {i, j, k};
                            if (g.interior(c, L::D)) continue;
                            for (int q = 0; q < QL; ++q) {
                                const size_t t = (size_t)q * hn() + hidx(i, j, k);
                                keep[t] = raw[t];
                            }
                        }
            }
        }
        recollide();
    }

    // NSAllen::macroscopic (NsAllen.cu:866-877) on the current f_post /
    // g_post: phi = sum g_post at interior nodes (the ghost ring is left as
    // it is), mu and grad phi from it, p = sum f_post and u from f_post with
    // u_old = the current u; then the pending collision.
    void macroscopic() override {
        materialise();
        HIP_CHECK(hipMemcpyAsync(phib[1 - phc], phib[phc], (size_t)g.nodes * sizeof(CT), hipMemcpyDeviceToDevice,
                                 stream));
        // pull buffer 0 in place, u[ucur] -> u[1-ucur], outflow planes -> prev[pcur]
        StepArgs<ST, CT> a = args(0, ucur, 1 - pcur, phib[1 - phc], nullptr);
        k_phase<L, ST, CT, SP, true, false><<<grid_interior(), block_shape(), 0, stream>>>(a, GhostJob<CT>{});
        launch_check("k_phase");
        k_fused<L, ST, CT, SP, true, MODE_FULL, false><<<grid_interior(), block_shape(), 0, stream>>>(a, 0u);
        launch_check("k_fused");
        const int tb = 256;
        k_export_post<L, ST, CT, SP, true><<<(unsigned)((g.nodes + tb - 1) / tb), tb, 0, stream>>>(a, true, scratch, p);
        launch_check("k_export_post");
        HIP_CHECK(hipStreamSynchronize(stream));
        phc = 1 - phc;
        ucur = 1 - ucur;
    }

    // phi, the phi the pulled buffer is encoded against, u and p of a handle
    // in the direct state (false otherwise).
    bool direct_state(std::vector<double>& phi_now, std::vector<double>& phi_ref, std::vector<double>& uu,
                      std::vector<double>& pp) override {
        if (!direct_old) return false;
        download_field(phib[phc], 1, phi_now);
        download_field(phib[1 - phc], 1, phi_ref);
        download_field(u[ucur], L::D, uu);
        download_field(p, 1, pp);
        return true;
    }

    // Restart into the direct state (the counterpart of direct_state, with
    // dev_f, dev_g, f_post, g_post from the checkpoint files); `raw_f` /
    // `raw_g` are the stored 16-bit bytes of the f_post / g_post buffer,
    // encoded against phi_ref unless `ref_zero`.
    void resume_direct(const std::vector<double>& f, const std::vector<double>& gg, const std::vector<double>& fpost,
                       const std::vector<double>& gpost, const std::vector<double>& phi_now,
                       const std::vector<double>& phi_ref, const std::vector<double>& uu,
                       const std::vector<double>& pp, bool ref_zero, const std::vector<unsigned char>* raw_f,
                       const std::vector<unsigned char>* raw_g) override {
        upload_field(phib[0], phi_ref, 1);
        upload_field(phib[1], phi_now, 1);
        phc = 1;
        ucur = 1;
        pcur = 0;
        upload_field(u[1], uu, L::D);
        upload_field(p, pp, 1);
        set_direct(fpost, gpost, f, gg);
        const size_t raw_bytes = hn() * QL * sizeof(ST);
        if (sizeof(ST) == 2 && raw_f && raw_g && raw_f->size() == raw_bytes && raw_g->size() == raw_bytes) {
            std::vector<unsigned char> df, dg;
            raw_to_dev(*raw_f, df);
            raw_to_dev(*raw_g, dg);
            HIP_CHECK(hipMemcpyAsync(fbuf[0], df.data(), df.size(), hipMemcpyHostToDevice, stream));
            HIP_CHECK(hipMemcpyAsync(gbuf[0], dg.data(), dg.size(), hipMemcpyHostToDevice, stream));
            HIP_CHECK(hipStreamSynchronize(stream));
            old_ref_zero = ref_zero;
        }
        initialized = true;
        recollide();
    }

    // Re-runs the collision that produced pop[cur] from the state the last
    // step left behind (used after a field was modified).
    void recollide() {
        const int old = 1 - cur;
        if (!direct_old) export_post(true);  // p from the current f_post
        // pull pop[old] with the u / outflow planes of that pull; collide with u[ucur]
        StepArgs<ST, CT> a = args(old, 1 - ucur, 1 - pcur, phib[phc], old_ref());
        a.ucol = u[ucur];
        a.uout = nullptr;
        if (direct_old) {
            k_fused<L, ST, CT, SP, true, MODE_COLLIDE, false><<<grid_interior(), block_shape(), 0, stream>>>(a, 0u);
            launch_check("k_fused");
        } else {
            if (shell_blocks()) launch_shell_pull(a);
            launch_fused_pull<MODE_COLLIDE>(a);
        }
        HIP_CHECK(hipStreamSynchronize(stream));
    }

    // Visits the inward populations of every outflow boundary node:
    // fn(side, q, node index, plane offset).
    template <class Fn>
    void for_outflow_entries(Fn fn) const {
        for (int s = 0; s < 2 * L::D; ++s) {
            if (cfg.bc_type[s] != CLIP_BC_FREE_CONVECT) continue;
            const int a = s / 2, inward = (s & 1) ? -1 : 1;
            for (int k = g.dmin[2]; k <= g.dmax[2]; ++k)
                for (int j = g.dmin[1]; j <= g.dmax[1]; ++j)
                    for (int i = g.dmin[0]; i <= g.dmax[0]; ++i) {
                        const int c[3] = {i, j, k};
                        if (c[a] != ((s & 1) ? g.dmax[a] : g.dmin[a])) continue;
                        for (int q = 0; q < QL; ++q)
                            if (L::e(q, a) == inward)
                                fn(s, q, (long long)hidx(i, j, k), (long long)q * plane_size(g, s) + plane_index(g, s, c));
                    }
        }
    }

    void prev_soa(bool is_f, std::vector<double>& out) override {
        get_field(is_f ? CLIP_FIELD_F_POST : CLIP_FIELD_G_POST, out);
        if (direct_old) return;
        const int set = 1 - pcur;  // the planes the last step's pull read
        for (int s = 0; s < 2 * L::D; ++s) {
            if (cfg.bc_type[s] != CLIP_BC_FREE_CONVECT) continue;
            std::vector<double> plane;
            download(is_f ? prevf[set][s] : prevg[set][s], (size_t)plane_size(g, s) * QL, plane);
            for_outflow_entries([&](int side, int q, long long node, long long off) {
                if (side == s) out[(size_t)q * hn() + node] = plane[off];
            });
        }
    }

    void phi_pair(std::vector<double>& now, std::vector<double>& before) override {
        download_field(phib[phc], 1, now);
        download_field(phib[1 - phc], 1, before);
    }

    bool raw_pulled(bool is_f, std::vector<unsigned char>& out) override {
        if constexpr (sizeof(ST) != 2) return false;
        if (!initialized) return false;
        const size_t bytes = (size_t)g.nodes * QL * sizeof(ST);
        std::vector<unsigned char> dev(bytes);
        HIP_CHECK(hipMemcpyAsync(dev.data(), is_f ? fbuf[1 - cur] : gbuf[1 - cur], bytes, hipMemcpyDeviceToHost,
                                 stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        raw_to_dense(dev, out);
        return true;
    }

    void resume(const std::vector<double>& f, const std::vector<double>& gg, const std::vector<double>& fpost,
                const std::vector<double>& gpost, const std::vector<double>* fprev, const std::vector<double>* gprev,
                const std::vector<double>* phi_now, const std::vector<double>* phi_before,
                const std::vector<double>* uold, const std::vector<unsigned char>* raw_f,
                const std::vector<unsigned char>* raw_g) override {
        const int tb = 256;
        if (phi_now && phi_before && uold) {
            // Exact restart: pop 0 = dev_f / dev_g (pulled) encoded relative to
            // the phi that wrote them, outflow planes from dev_f_prev /
            // dev_g_prev, u_old and both phi fields from the extension; then the
            // step kernels recompute macroscopic() and the pending collision.
            upload_field(phib[0], *phi_before, 1);
            upload_field(phib[1], *phi_now, 1);
            const size_t raw_bytes = hn() * QL * sizeof(ST);
            if (sizeof(ST) == 2 && raw_f && raw_g && raw_f->size() == raw_bytes && raw_g->size() == raw_bytes) {
                // the stored 16-bit bytes themselves: no re-encoding
                std::vector<unsigned char> df, dg;
                raw_to_dev(*raw_f, df);
                raw_to_dev(*raw_g, dg);
                for (int b = 0; b < 2; ++b) {
                    HIP_CHECK(hipMemcpyAsync(fbuf[b], df.data(), df.size(), hipMemcpyHostToDevice, stream));
                    HIP_CHECK(hipMemcpyAsync(gbuf[b], dg.data(), dg.size(), hipMemcpyHostToDevice, stream));
                }
                HIP_CHECK(hipStreamSynchronize(stream));
            } else {
                upload_pop(fbuf[0], f, nullptr);
                upload_pop(gbuf[0], gg, phib[0]);
                upload_pop(fbuf[1], f, nullptr);
                upload_pop(gbuf[1], gg, phib[0]);
            }
            upload_field(u[0], *uold, L::D);
            for (int s = 0; s < 2 * L::D; ++s) {
                if (cfg.bc_type[s] != CLIP_BC_FREE_CONVECT) continue;
                std::vector<double> pf((size_t)plane_size(g, s) * QL, 0.0), pg(pf.size(), 0.0);
                for_outflow_entries([&](int side, int q, long long node, long long off) {
                    if (side != s) return;
                    pf[off] = fprev ? (*fprev)[(size_t)q * hn() + node] : fpost[(size_t)q * hn() + node];
                    pg[off] = gprev ? (*gprev)[(size_t)q * hn() + node] : gpost[(size_t)q * hn() + node];
                });
                upload(prevf[0][s], pf);
                upload(prevg[0][s], pg);
            }
            StepArgs<ST, CT> a = args(0, 0, 0, phib[1], phib[0]);
            launch_phase_pull(a);
            launch_fused_pull<MODE_FULL>(a);
            HIP_CHECK(hipStreamSynchronize(stream));
            cur = 1;
            ucur = 1;
            pcur = 1;
            phc = 1;
            direct_old = false;
            old_ref_zero = false;
            direct_f_interior.clear();
            direct_g_interior.clear();
            initialized = true;
            return;
        }
        // Reference restart (main.cu:67-71): pop 0 holds dev_f / dev_g with the
        // interior replaced by f_post / g_post (read directly, absolute),
        // pop 1 dev_f / dev_g; macroscopic() with u_old = 0, then the collision.
        std::vector<double> a0 = f, b0 = gg;
        for (int k = g.dmin[2]; k <= g.dmax[2]; ++k)
            for (int j = g.dmin[1]; j <= g.dmax[1]; ++j)
                for (int i = g.dmin[0]; i <= g.dmax[0]; ++i)
                    for (int q = 0; q < QL; ++q) {
                        const size_t t = (size_t)q * hn() + hidx(i, j, k);
                        a0[t] = fpost[t];
                        b0[t] = gpost[t];
                    }
        upload_pop(fbuf[0], a0, nullptr);
        upload_pop(gbuf[0], b0, nullptr);
        upload_pop(fbuf[1], f, nullptr);
        upload_pop(gbuf[1], gg, nullptr);
        std::vector<double> zero(hn() * L::D, 0.0);
        upload_field(u[0], zero, L::D);
        StepArgs<ST, CT> a = args(0, 0, 1, phib[1], nullptr);
        k_phase<L, ST, CT, SP, true, false><<<grid_interior(), block_shape(), 0, stream>>>(a, GhostJob<CT>{});
        launch_check("k_phase");
        // ghost ring from the resumed phi itself
        HIP_CHECK(hipMemcpyAsync(phib[0], phib[1], (size_t)g.nodes * sizeof(CT), hipMemcpyDeviceToDevice, stream));
        ghost_refresh(phib[0], phib[1]);
        k_fused<L, ST, CT, SP, true, MODE_FULL, false><<<grid_interior(), block_shape(), 0, stream>>>(a, 0u);
        launch_check("k_fused");
        k_export_post<L, ST, CT, SP, true><<<(unsigned)((g.nodes + tb - 1) / tb), tb, 0, stream>>>(a, true, scratch, p);
        launch_check("k_export_post");
        HIP_CHECK(hipStreamSynchronize(stream));
        cur = 1;
        ucur = 1;
        pcur = 0;
        phc = 1;
        direct_old = true;
        old_ref_zero = true;
        direct_f_interior = f;
        direct_g_interior = gg;
        initialized = true;
    }

    // ------------------------------------------------------------- adjoint
    // Reverse mode over nsteps steps from a given interior phi0.  The forward
    // keeps a state every `every` steps; the sweep recomputes each segment.
    // With an inner interval (0 < inner < every) the recomputation keeps a
    // state every `inner` steps of the segment and recomputes each of those
    // sub-segments once more (two checkpoint levels: memory every/inner +
    // inner states instead of every, one more forward pass).
    std::vector<void*> adj_allocs;
    size_t adj_capacity = 0;  // states in the pool
    size_t adj_bytes = 0;
    CT* adj_pool = nullptr;
    CT *lF = nullptr, *lG = nullptr, *lF2 = nullptr, *lG2 = nullptr, *lU = nullptr, *lPhi = nullptr, *lfp = nullptr,
       *lgp = nullptr,
       *lst = nullptr, *lphin = nullptr;
    StAdj<CT>* lsv = nullptr;  // adjoint of grad phi, lap phi per node (k_adj_node -> k_adj_phi)
    double* adj_io = nullptr;  // FP64 staging of the interior arrays (interior nodes)
    CT* lP[2] = {nullptr, nullptr};  // adjoint of the outflow planes (f sides, then g sides), ping-pong
    int lpi = 0;                     // lP[lpi] holds the adjoint of the planes of the state after the step
    double *adj_part = nullptr, *adj_acc = nullptr, *in_accf = nullptr, *in_accg = nullptr;
    int adj_T = 0, adj_K = 0, adj_inner = 0, adj_nsnap = 0, adj_nmid = 0, adj_leaf = 0;
    double adj_nfloor = 0.0;
    bool adj_ready = false, adj_grad_ready = false;
    std::vector<double> adj_inlet_grad;
    std::vector<CT> adj_phi0;  // phi of the initial state on every node (device layout)
    // host-built transposes (once per handle)
    bool maps_ready = false;
    AdjRows rows_pop{}, rows_prev{}, rows_inlet{}, rows_u{}, rows_ring{}, rows_init{};
    std::vector<long long> inlet_rows;  // population index (q*nodes + ghost node) of each inlet row
    long long poff[6] = {};
    long long prev_elems = 0;           // entries of one plane set over all free-convect sides

    // a state rounded up to whole 128-byte lines, so that every part of every
    // pool slot keeps the row alignment of alloc_grid
    size_t state_elems() const {
        const size_t line = 128 / sizeof(CT);
        return ((size_t)g.nodes * (2 * QL + 1 + L::D) + 2 * (size_t)prev_elems + line - 1) / line * line;
    }
    // pool: kept states 0, K, 2K, ... | the final state | inner-level states | leaf states
    CT* slot(size_t k) { return adj_pool + k * state_elems(); }
    CT* snap_ptr(int k) { return slot((size_t)k); }
    CT* final_ptr() { return slot((size_t)adj_nsnap); }
    CT* mid_ptr(int j) { return slot((size_t)(adj_nsnap + 1 + j)); }
    CT* leaf_ptr(int j) { return slot((size_t)(adj_nsnap + 1 + adj_nmid + j)); }
    size_t pool_states() const {
        return (size_t)adj_nsnap + 1 + adj_nmid + (size_t)std::max(adj_leaf - 1, 2);
    }
    // parts of a kept state: f*, g*, phi (with its ghost ring), u, outflow planes
    CT* st_f(CT* st) const { return st; }
    CT* st_g(CT* st) const { return st + (size_t)g.nodes * QL; }
    CT* st_phi(CT* st) const { return st + (size_t)g.nodes * 2 * QL; }
    CT* st_u(CT* st) const { return st + (size_t)g.nodes * (2 * QL + 1); }
    CT* st_pf(CT* st) const { return st + (size_t)g.nodes * (2 * QL + 1 + L::D); }
    CT* st_pg(CT* st) const { return st_pf(st) + prev_elems; }

    // One forward step from kept state `a` into kept state `b` (no copies).
    void step_into(CT* a, CT* b) {
        if constexpr (kPlain) {
            StepArgs<ST, CT> s_ = args(cur, ucur, pcur, st_phi(b), st_phi(a));
            s_.fin = (const ST*)st_f(a);
            s_.gin = (const ST*)st_g(a);
            s_.fout = (ST*)st_f(b);
            s_.gout = (ST*)st_g(b);
            s_.uin = st_u(a);
            s_.ufc = st_u(a);
            s_.ucol = st_u(a);
            s_.uout = st_u(b);
            for (int s = 0; s < 6; ++s) {
                const bool fc = s < 2 * L::D && cfg.bc_type[s] == CLIP_BC_FREE_CONVECT;
                s_.prev_f_in[s] = fc ? st_pf(a) + poff[s] : nullptr;
                s_.prev_g_in[s] = fc ? st_pg(a) + poff[s] : nullptr;
                s_.prev_f_out[s] = fc ? st_pf(b) + poff[s] : nullptr;
                s_.prev_g_out[s] = fc ? st_pg(b) + poff[s] : nullptr;
            }
            if (tile_step()) {
                launch_step_tile(s_, st_phi(a), 1);
                return;
            }
            const GhostJob<CT> gj = ghost_job(st_phi(a), st_phi(b));
            if (!gj.total) ghost_refresh(st_phi(a), st_phi(b));
            launch_phase_pull(s_, gj);
            launch_fused_pull<MODE_FULL>(s_);
        }
    }
    long long interior_count() const {
        long long t = 1;
        for (int a = 0; a < L::D; ++a) t *= g.dmax[a] - g.dmin[a] + 1;
        return t;
    }
    // Element of interior node (i, j, k) in a host interior array, x fastest
    // (xslow = 0) or x slowest (xslow = 1); k_adj_from_interior's order.
    long long interior_index(int i, int j, int k, int xslow) const {
        const long long n0 = g.dmax[0] - g.dmin[0] + 1, n1 = g.dmax[1] - g.dmin[1] + 1, n2 = g.dmax[2] - g.dmin[2] + 1;
        const long long li = i - g.dmin[0], lj = j - g.dmin[1], lk = k - g.dmin[2];
        return xslow ? (li * n1 + lj) * n2 + lk : li + n0 * (lj + n1 * lk);
    }
    // Host interior array -> interior nodes of a node array (one transfer
    // into the FP64 staging buffer, scattered on the device).
    void interior_in(CT* dst, const double* src, int xslow) {
        const long long n = interior_count();
        HIP_CHECK(hipMemcpyAsync(adj_io, src, (size_t)n * sizeof(double), hipMemcpyHostToDevice, stream));
        k_adj_from_interior<CT><<<(unsigned)((n + ADJ_TB - 1) / ADJ_TB), ADJ_TB, 0, stream>>>(dst, adj_io, g, n, xslow);
        launch_check("k_adj_from_interior");
        HIP_CHECK(hipStreamSynchronize(stream));
    }
    // a (+ b where non-null) at the interior nodes -> host interior array.
    void interior_out(double* dst, const CT* a, const CT* b, int xslow) {
        const long long n = interior_count();
        k_adj_to_interior<CT><<<(unsigned)((n + ADJ_TB - 1) / ADJ_TB), ADJ_TB, 0, stream>>>(adj_io, a, b, g, n, xslow);
        launch_check("k_adj_to_interior");
        HIP_CHECK(hipMemcpyAsync(dst, adj_io, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
    }
    void adj_check() const {
        if constexpr (!std::is_same<ST, CT>::value)
            throw ConfigError("the adjoint runs on F64 or F32 handles");
        for (int s = 0; s < 2 * L::D; ++s)
            if (cfg.bc_type[s] == CLIP_BC_SLIP_WALL && !bc.wall_pass)
                throw ConfigError("the adjoint does not support a slip side without a wall side "
                                  "(unless slip sides reflect specularly)");
    }
    template <class T>
    T* adj_alloc(size_t count) {
        void* ptr = nullptr;
        if (hipMalloc(&ptr, count * sizeof(T)) != hipSuccess) {
            (void)hipGetLastError();
            throw std::bad_alloc();
        }
        adj_allocs.push_back(ptr);
        adj_bytes += count * sizeof(T);
        bytes += count * sizeof(T);
        return (T*)ptr;
    }
    // Node array of the reverse sweep with the row alignment of alloc_grid.
    template <class T>
    T* adj_alloc_grid(size_t count) {
        const size_t lead = row_align ? 128 / sizeof(T) - 1 : 0;
        return adj_alloc<T>(count + lead + 1) + lead;
    }
    void adj_free() {
        for (void* p2 : adj_allocs) (void)hipFree(p2);
        adj_allocs.clear();
        bytes -= adj_bytes;
        adj_bytes = 0;
        adj_capacity = 0;
        adj_pool = nullptr;
        maps_ready = false;
    }
    void save_state(CT* dst) {
        static_assert(kPlain, "plain storage only");
        const size_t nq = (size_t)g.nodes * QL;
        HIP_CHECK(hipMemcpyAsync(dst, fbuf[cur], nq * sizeof(CT), hipMemcpyDeviceToDevice, stream));
        HIP_CHECK(hipMemcpyAsync(dst + nq, gbuf[cur], nq * sizeof(CT), hipMemcpyDeviceToDevice, stream));
        HIP_CHECK(hipMemcpyAsync(st_phi(dst), phib[phc], (size_t)g.nodes * sizeof(CT), hipMemcpyDeviceToDevice,
                                 stream));
        HIP_CHECK(hipMemcpyAsync(st_u(dst), u[ucur], (size_t)g.nodes * L::D * sizeof(CT), hipMemcpyDeviceToDevice,
                                 stream));
        for (int s = 0; s < 2 * L::D; ++s) {
            if (cfg.bc_type[s] != CLIP_BC_FREE_CONVECT) continue;
            const size_t n = (size_t)plane_size(g, s) * QL * sizeof(CT);
            HIP_CHECK(hipMemcpyAsync(st_pf(dst) + poff[s], prevf[pcur][s], n, hipMemcpyDeviceToDevice, stream));
            HIP_CHECK(hipMemcpyAsync(st_pg(dst) + poff[s], prevg[pcur][s], n, hipMemcpyDeviceToDevice, stream));
        }
    }

    static constexpr bool kPlain = std::is_same<ST, CT>::value;
    // Whether the bulk collision launch carries the boundary layer: not for
    // FP64 D3Q19, whose boundary-layer path needs 222 VGPRs against the
    // bulk's 164 (the carrying kernel would drop from 3 to 2 waves per SIMD).
    static constexpr bool kCarryShell = !(std::is_same<CT, double>::value && L::Q == 19);

    bool shell_host(const int c[3]) const {
        bool s = false;
        for (int a = 0; a < L::D; ++a) s |= !bc.both_periodic[a] && (c[a] == g.dmin[a] || c[a] == g.dmax[a]);
        return s;
    }
    std::vector<std::vector<unsigned char>> inlet_flags_host() const { return inlet_flags(cfg, g, L::D); }
    // (target, entry) pairs -> row lists in device memory (rows by target,
    // entries by reader: a fixed order)
    AdjRows upload_rows(std::vector<std::pair<long long, AdjEntry>>& v, std::vector<long long>* targets = nullptr) {
        std::sort(v.begin(), v.end(), [](const std::pair<long long, AdjEntry>& x, const std::pair<long long, AdjEntry>& y) {
            return x.first != y.first ? x.first < y.first : x.second.rd < y.second.rd;
        });
        std::vector<long long> tg;
        std::vector<int> st;
        std::vector<AdjEntry> en;
        for (size_t i = 0; i < v.size(); ++i) {
            if (i == 0 || v[i].first != v[i - 1].first) {
                tg.push_back(v[i].first);
                st.push_back((int)en.size());
            }
            en.push_back(v[i].second);
        }
        st.push_back((int)en.size());
        if (targets) *targets = tg;
        AdjRows r{};
        r.rows = (int)tg.size();
        if (!r.rows) return r;
        long long* dt = adj_alloc<long long>(tg.size());
        int* ds = adj_alloc<int>(st.size());
        AdjEntry* de = adj_alloc<AdjEntry>(en.size());
        HIP_CHECK(hipMemcpy(dt, tg.data(), tg.size() * sizeof(long long), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(ds, st.data(), st.size() * sizeof(int), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(de, en.data(), en.size() * sizeof(AdjEntry), hipMemcpyHostToDevice));
        r.target = dt;
        r.start = ds;
        r.e = de;
        return r;
    }
    // Transposes of the boundary-layer pulls, the phi ring copies and the
    // initial ring (SymPuller, ring_sources).
    void maps_build() {
        if (maps_ready) return;
        const auto vf = inlet_flags_host();
        SymPuller<L> sp{g, bc, vf};
        std::vector<std::pair<long long, AdjEntry>> pop, prv, inl, uu;
        std::vector<PullTerm> terms;
        for (int k = g.dmin[2]; k <= g.dmax[2]; ++k)
            for (int j = g.dmin[1]; j <= g.dmax[1]; ++j)
                for (int i = g.dmin[0]; i <= g.dmax[0]; ++i) {
                    const int c[3] = {i, j, k};
                    if (!shell_host(c)) continue;
                    const long long x = g.idx(i, j, k);
                    for (int q = 0; q < QL; ++q) {
                        terms.clear();
                        sp.lvl3(c, q, terms);
                        const long long rd = (long long)q * g.nodes + x;
                        const PullTerm* X = nullptr;
                        const PullTerm* P = nullptr;
                        for (const auto& t : terms) {
                            AdjEntry e{rd, t.fc >= 0 ? (long long)(t.fc / 2) * g.nodes + t.nb : -1, -1, -1, t.branch, 0};
                            if (t.kind == 0) pop.push_back({t.idx, e});
                            else if (t.kind == 1) prv.push_back({poff[t.side] + t.idx, e});
                            else if (t.kind == 2 || t.kind == 3) inl.push_back({t.idx, e});
                            if (t.fc >= 0 && t.kind == 1) P = &t;
                            if (t.fc >= 0 && t.kind != 1) X = &t;
                        }
                        if (P) {
                            if (!X) throw ConfigError("adjoint: free-convect pull without a neighbour value");
                            uu.push_back({(long long)(P->fc / 2) * g.nodes + P->nb,
                                          AdjEntry{rd, -1, X->idx, P->idx, 0, P->side}});
                        }
                    }
                }
        rows_pop = upload_rows(pop);
        rows_prev = upload_rows(prv);
        rows_inlet = upload_rows(inl, &inlet_rows);
        rows_u = upload_rows(uu);
        // ring copies of one step, and of the initial phi (ghosts wrapped
        // along periodic axes copy phi0; others are constants)
        const std::vector<long long> src = ring_sources(g, bc, L::D, vf);
        std::vector<std::pair<long long, AdjEntry>> ring, init;
        for (long long n = 0; n < g.nodes; ++n)
            if (src[n] >= 0) ring.push_back({src[n], AdjEntry{n, -1, -1, -1, 0, 0}});
        for (int k = 0; k < g.ext[2]; ++k)
            for (int j = 0; j < g.ext[1]; ++j)
                for (int i = 0; i < g.ext[0]; ++i) {
                    const int c[3] = {i, j, k};
                    if (g.interior(c, L::D)) continue;
                    int w[3] = {i, j, k};
                    bool wrapped = true;
                    for (int a = 0; a < L::D; ++a) {
                        if (w[a] >= g.dmin[a] && w[a] <= g.dmax[a]) continue;
                        if (!bc.periodic[a]) wrapped = false;
                        const int span = g.dmax[a] - g.dmin[a] + 1;
                        w[a] += (w[a] < g.dmin[a]) ? span : -span;
                    }
                    if (wrapped) init.push_back({g.idx(w[0], w[1], w[2]), AdjEntry{g.idx(i, j, k), -1, -1, -1, 0, 0}});
                }
        rows_ring = upload_rows(ring);
        rows_init = upload_rows(init);
        maps_ready = true;
    }

    void adjoint_forward(const double* phi0, int xslow, int nsteps, int every, double* out) override {
        adj_check();
        if constexpr (kPlain) adjoint_forward_impl(phi0, xslow, nsteps, every, out);
    }
    void adjoint_forward_impl(const double* phi0, int xslow, int nsteps, int every, double* out) {
        if (nsteps < 1 || every < 1) throw ConfigError("adjoint: nsteps and checkpoint interval must be >= 1");
        // phi of the ghost cells: a ghost cell copies phi0 at its wrapped
        // position when every axis it lies outside of is periodic, else it
        // keeps the case's initial phase field (a constant); the interior
        // comes from phi0 on the device (interior_in below)
        std::vector<CT> h((size_t)g.nodes, CT(0));
        for (int k = 0; k < g.ext[2]; ++k)
            for (int j = 0; j < g.ext[1]; ++j) {
                const bool row_in = j >= g.dmin[1] && j <= g.dmax[1] &&
                                    (L::D == 2 || (k >= g.dmin[2] && k <= g.dmax[2]));
                for (int i = 0; i < g.ext[0]; ++i) {
                    if (row_in && i == g.dmin[0]) {
                        i = g.dmax[0];
                        continue;
                    }
                    int w[3] = {i, j, k};
                    bool wrapped = true;
                    for (int a = 0; a < L::D; ++a) {
                        if (w[a] >= g.dmin[a] && w[a] <= g.dmax[a]) continue;
                        if (!bc.periodic[a]) wrapped = false;
                        const int span = g.dmax[a] - g.dmin[a] + 1;
                        w[a] += (w[a] < g.dmin[a]) ? span : -span;
                    }
                    h[g.idx(i, j, k)] =
                        wrapped ? (CT)phi0[interior_index(w[0], w[1], w[2], xslow)] : (CT)initial_phi(cfg, i, j, k);
                }
            }
        prev_elems = 0;
        for (int s = 0; s < 6; ++s) {
            poff[s] = prev_elems;
            if (s < 2 * L::D && cfg.bc_type[s] == CLIP_BC_FREE_CONVECT) prev_elems += plane_size(g, s) * QL;
        }
        adj_T = nsteps;
        adj_K = every;
        adj_nsnap = (nsteps + every - 1) / every;
        if (adj_inner > 0 && adj_inner < every) {
            adj_leaf = adj_inner;
            adj_nmid = (every + adj_inner - 1) / adj_inner - 1;
        } else {
            adj_leaf = every;
            adj_nmid = 0;
        }
        const size_t need = pool_states();
        if (need > adj_capacity || !adj_pool) {
            adj_free();
            adj_pool = adj_alloc_grid<CT>(need * state_elems());
            adj_capacity = need;
            const size_t nq = (size_t)g.nodes * QL;
            lF = adj_alloc_grid<CT>(nq);
            lG = adj_alloc_grid<CT>(nq);
            lF2 = adj_alloc_grid<CT>(nq);
            lG2 = adj_alloc_grid<CT>(nq);
            lfp = adj_alloc_grid<CT>(nq);
            lgp = adj_alloc_grid<CT>(nq);
            lU = adj_alloc_grid<CT>((size_t)g.nodes * L::D);
            lPhi = adj_alloc_grid<CT>((size_t)g.nodes);
            lphin = adj_alloc_grid<CT>((size_t)g.nodes);
            lst = adj_alloc_grid<CT>((size_t)g.nodes);
            // records of 4 CT, 16-byte aligned or more; node x = 1 of a row on a
            // 128-byte line (the row stride is a multiple of 128 bytes of CT)
            lsv = adj_alloc<StAdj<CT>>((size_t)g.nodes + 8) + (row_align ? 128 / sizeof(StAdj<CT>) - 1 : 0);
            adj_io = adj_alloc<double>((size_t)interior_count());
            if (prev_elems) {
                lP[0] = adj_alloc<CT>(2 * (size_t)prev_elems);
                lP[1] = adj_alloc<CT>(2 * (size_t)prev_elems);
            } else {
                lP[0] = lP[1] = nullptr;
            }
            const long long blocks = (interior_count() + ADJ_TB - 1) / ADJ_TB;
            adj_part = adj_alloc<double>((size_t)blocks * 4);
            adj_acc = adj_alloc<double>(4);
            maps_build();
            in_accf = adj_alloc<double>(std::max<size_t>(1, inlet_rows.size()));
            in_accg = adj_alloc<double>(std::max<size_t>(1, inlet_rows.size()));
        }
        HIP_CHECK(hipMemcpyAsync(phib[0], h.data(), h.size() * sizeof(CT), hipMemcpyHostToDevice, stream));
        interior_in(phib[0], phi0, xslow);
        HIP_CHECK(hipMemcpyAsync(phib[1], phib[0], h.size() * sizeof(CT), hipMemcpyDeviceToDevice, stream));
        init_from_phib();
        adj_phi0 = std::move(h);
        save_state(snap_ptr(0));
        // the recomputation slots (inner level, leaves) start as a copy of the
        // initial state: cells no step writes (ghost populations, constant
        // ring cells) hold their values.  The kept states and the final one
        // are whole copies of the handle's state (save_state).
        for (size_t k = (size_t)adj_nsnap + 1; k < need; ++k)
            HIP_CHECK(hipMemcpyAsync(slot(k), slot(0), state_elems() * sizeof(CT), hipMemcpyDeviceToDevice, stream));
        for (int n = 1; n <= nsteps; ++n) {
            step(1);
            if (n % every == 0 && n < nsteps) save_state(snap_ptr(n / every));
        }
        save_state(final_ptr());
        interior_out(out, phib[phc], nullptr, xslow);
        adj_ready = true;
        adj_grad_ready = false;
        step_count = nsteps;
    }

    // Arguments of the reverse of the step from kept state `st` to the state
    // `after` (whose phi is the step's stencil input with its lagged ring).
    AdjArgs<CT> adj_args(CT* st, CT* after) {
        AdjArgs<CT> A;
        A.g = g;
        A.bc = bc;
        A.ph = ph;
        A.F = st_f(st);
        A.G = st_g(st);
        A.phin = st_phi(after);
        A.uo = st_u(st);
        A.lF = lF;
        A.lG = lG;
        A.lFo = lF2;
        A.lGo = lG2;
        A.lU = lU;
        A.lPhi = lPhi;
        A.lfp = lfp;
        A.lgp = lgp;
        A.lst = lst;
        A.lsv = lsv;
        A.lphin = lphin;
        A.part = adj_part;
        A.sigma = (CT)cfg.sigma;
        A.half_length = (CT)(cfg.reference_length / 2.0);
        A.nfloor = (CT)adj_nfloor;
        A.pf = prev_elems ? st_pf(st) : nullptr;
        A.pg = prev_elems ? st_pg(st) : nullptr;
        A.lPf_in = lP[lpi];
        A.lPg_in = lP[lpi] ? lP[lpi] + prev_elems : nullptr;
        A.lPf_out = lP[1 - lpi];
        A.lPg_out = lP[1 - lpi] ? lP[1 - lpi] + prev_elems : nullptr;
        for (int s = 0; s < 6; ++s) A.poff[s] = poff[s];
        return A;
    }

    void fill(CT* x, long long n) {
        if (n <= 0) return;
        k_adj_fill<CT><<<(unsigned)((n + ADJ_TB - 1) / ADJ_TB), ADJ_TB, 0, stream>>>(x, n, CT(0));
        launch_check("k_adj_fill");
    }
    template <class Fn>
    void launch_rows(const AdjRows& r, Fn fn) {
        if (r.rows) fn((unsigned)((r.rows + ADJ_TB - 1) / ADJ_TB));
    }

    // Reverse of the step from kept state `st` into `after`: lF / lG (the
    // adjoint of the later state's populations) -> lF2 / lG2 (that of st's),
    // then the two swap.  The partial sums of k_adj_node's parameter
    // adjoints are folded into the accumulator by an extra block of the
    // k_adj_phi launch.
    void adjoint_step(CT* st, CT* after) {
        const AdjArgs<CT> A = adj_args(st, after);
        const unsigned blocks = (unsigned)((interior_count() + ADJ_TB - 1) / ADJ_TB);
        const unsigned nb = (unsigned)((g.nodes + ADJ_TB - 1) / ADJ_TB);
        if (shell_blocks()) {
            k_adj_shell_post<L, CT><<<dim3(shell_blocks(), QL), ADJ_TB, 0, stream>>>(A, make_shell<L>(g, bc));
            launch_check("k_adj_shell_post");
        }
        k_adj_node<L, CT><<<blocks, ADJ_TB, 0, stream>>>(A);
        launch_check("k_adj_node");
        if (!shell_nodes()) {
            k_adj_phi_periodic<L, CT, true><<<blocks + 1, ADJ_TB, 0, stream>>>(A, adj_acc, (int)blocks);
            launch_check("k_adj_phi");
        } else {
            k_adj_phi<L, CT, true><<<nb + 1, ADJ_TB, 0, stream>>>(A, lphin, adj_acc, (int)blocks);
            launch_check("k_adj_phi");
            fill(lPhi, g.nodes);
            launch_rows(rows_ring,
                        [&](unsigned b) { k_adj_ring<CT><<<b, ADJ_TB, 0, stream>>>(rows_ring, lphin, lPhi); });
        }
        launch_rows(rows_pop, [&](unsigned b) { k_adj_csr_pop<CT><<<b, ADJ_TB, 0, stream>>>(A, rows_pop); });
        if (prev_elems) {
            fill(lP[1 - lpi], 2 * prev_elems);
            launch_rows(rows_prev, [&](unsigned b) { k_adj_csr_prev<CT><<<b, ADJ_TB, 0, stream>>>(A, rows_prev); });
        }
        launch_rows(rows_inlet,
                    [&](unsigned b) { k_adj_csr_inlet<CT><<<b, ADJ_TB, 0, stream>>>(A, rows_inlet, in_accf, in_accg); });
        launch_rows(rows_u, [&](unsigned b) { k_adj_csr_u<CT><<<b, ADJ_TB, 0, stream>>>(A, rows_u); });
        launch_check("k_adj_csr");
        lpi = 1 - lpi;
        std::swap(lF, lF2);
        std::swap(lG, lG2);
    }

    // Steps s0 .. s1-1 in reverse, from kept state `sa` (state s0); `after`
    // is state s1.  Leaf level: every state recomputed into the leaf slots.
    void reverse_leaf(CT* sa, int s0, int s1, CT* after) {
        auto state = [&](int n) -> CT* { return n == s0 ? sa : leaf_ptr(n - s0 - 1); };
        for (int n = s0 + 1; n < s1; ++n) step_into(state(n - 1), state(n));
        for (int n = s1 - 1; n >= s0; --n) adjoint_step(state(n), n + 1 < s1 ? state(n + 1) : after);
    }
    void reverse_segment(CT* sa, int s0, int s1, CT* after) {
        if (adj_nmid == 0 || s1 - s0 <= adj_leaf) return reverse_leaf(sa, s0, s1, after);
        // inner level: recompute the segment keeping every adj_leaf-th state
        // (leaf slots 0 / 1 carry the states in between)
        CT* x = sa;
        for (int n = s0 + 1; n < s1; ++n) {
            const int r = n - s0;
            CT* dst = (r % adj_leaf == 0) ? mid_ptr(r / adj_leaf - 1) : leaf_ptr(n & 1);
            step_into(x, dst);
            x = dst;
        }
        const int chunks = (s1 - s0 + adj_leaf - 1) / adj_leaf;
        for (int c = chunks - 1; c >= 0; --c) {
            const int c0 = s0 + c * adj_leaf, c1 = std::min(s1, c0 + adj_leaf);
            reverse_leaf(c == 0 ? sa : mid_ptr(c - 1), c0, c1, c1 < s1 ? mid_ptr(c) : after);
        }
    }

    void adjoint_options(double normal_floor) override {
        if (!(normal_floor >= 0.0)) throw ConfigError("adjoint: normal floor must be >= 0");
        adj_nfloor = normal_floor;
    }
    void adjoint_nesting(int inner) override {
        if (inner < 0) throw ConfigError("adjoint: the inner checkpoint interval must be >= 0");
        adj_inner = inner;
    }

    void adjoint_backward(const double* gout, int xslow, double* gphi0, double* gsigma, double* gmob) override {
        adj_check();
        if constexpr (kPlain) adjoint_backward_impl(gout, xslow, gphi0, gsigma, gmob);
    }
    void adjoint_backward_impl(const double* gout, int xslow, double* gphi0, double* gsigma, double* gmob) {
        if (!adj_ready) throw ConfigError("adjoint: call the adjoint forward first");
        const size_t nq = (size_t)g.nodes * QL;
        fill(lF, (long long)nq);
        fill(lG, (long long)nq);
        fill(lF2, (long long)nq);
        fill(lG2, (long long)nq);
        fill(lU, g.nodes * L::D);
        fill(lst, g.nodes);
        fill((CT*)lsv, g.nodes * 4);
        fill(lphin, g.nodes);
        lpi = 0;
        if (prev_elems) {
            fill(lP[0], 2 * prev_elems);
            fill(lP[1], 2 * prev_elems);
        }
        HIP_CHECK(hipMemsetAsync(adj_acc, 0, 4 * sizeof(double), stream));
        HIP_CHECK(hipMemsetAsync(in_accf, 0, std::max<size_t>(1, inlet_rows.size()) * sizeof(double), stream));
        HIP_CHECK(hipMemsetAsync(in_accg, 0, std::max<size_t>(1, inlet_rows.size()) * sizeof(double), stream));
        fill(lPhi, g.nodes);
        interior_in(lPhi, gout, xslow);
        for (int k = adj_nsnap - 1; k >= 0; --k) {
            const int s0 = k * adj_K;
            const int s1 = std::min(adj_T, s0 + adj_K);
            reverse_segment(snap_ptr(k), s0, s1, s1 < adj_T ? snap_ptr(k + 1) : final_ptr());
        }
        // initial state and first collision
        const AdjArgs<CT> A = adj_args(snap_ptr(0), snap_ptr(0));
        const unsigned blocks = (unsigned)((interior_count() + ADJ_TB - 1) / ADJ_TB);
        const unsigned nb = (unsigned)((g.nodes + ADJ_TB - 1) / ADJ_TB);
        k_adj_init<L, CT><<<blocks, ADJ_TB, 0, stream>>>(A);
        launch_check("k_adj_init");
        k_adj_accumulate<CT><<<1, ADJ_TB, 0, stream>>>(adj_part, (int)blocks, adj_acc);
        launch_check("k_adj_accumulate");
        if (!shell_nodes()) {
            k_adj_phi_periodic<L, CT, false><<<blocks, ADJ_TB, 0, stream>>>(A, nullptr, 0);
            launch_check("k_adj_phi");
            interior_out(gphi0, lphin, nullptr, xslow);
        } else {
            k_adj_phi<L, CT, false><<<nb, ADJ_TB, 0, stream>>>(A, lphin, nullptr, 0);
            launch_check("k_adj_phi");
            fill(lPhi, g.nodes);
            launch_rows(rows_init,
                        [&](unsigned b) { k_adj_ring<CT><<<b, ADJ_TB, 0, stream>>>(rows_init, lphin, lPhi); });
            interior_out(gphi0, lphin, lPhi, xslow);
        }
        double acc[4];
        HIP_CHECK(hipMemcpyAsync(acc, adj_acc, sizeof(acc), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        // beta = 8 sigma / W, kappa = 1.5 sigma W (InputData.cu:76-114)
        const double gs_ghost = ghost_gradients();
        if (gsigma)
            *gsigma = acc[0] * 8.0 / cfg.interface_width + acc[1] * 1.5 * cfg.interface_width + acc[3] + gs_ghost;
        if (gmob) *gmob = acc[2];
        // leave the handle at the final state of the forward run
        adj_ready = false;
        adj_grad_ready = true;
    }

    // Gradients through the ghost values the pulls read (accumulated over the
    // sweep per row of the ghost list).  Inside an inlet disc: dL/dA_m of the
    // inlet modulation (below).  Elsewhere (do-nothing sides): the initial
    // state's f = p w with the Laplace pressure p = +-phi sigma / (L/2) /
    // (rho/3) of the drop / bubble cases (KernelInitializeDistributions,
    // NsAllen.cu:155-168), whose sigma-derivative is returned; g = phi w there
    // (the normal is 0 at ghosts) depends on no parameter.
    //
    // dL/dA_m of the inlet modulation u_in = value (1 + sum_m A_m cos((m+1) theta
    // + phase_m)) from the adjoints of the inlet ghost values: those values
    // (JetBoundary, Solver.cu:578-619) are, with G = Gamma(U) and the ghost
    // cell's constant populations f, g and phi,
    //   f_a = f_b + G_a - G_b - (1/weight) e_a,t . M_t,  M_t = sum_{q>=1, e_n=0} e_q,t (f_q - G_q)
    //   g_a = phi (G_a + w_a) - (g_b - phi (G_b + w_b)) - [a not normal] N / weight,
    //         N = sum_{q>=1, e_n=0} (g_q - phi (G_q + w_q))
    // with b = opp(a), so their U-derivatives need only dG_q/dU = w_q (3 e_q +
    // 9 (e_q.U) e_q - 3 U).
    double ghost_gradients() {
        adj_inlet_grad.assign(CLIP_MAX_INLET_MODES, 0.0);
        if (inlet_rows.empty()) return 0.0;
        std::vector<double> af(inlet_rows.size()), ag(inlet_rows.size());
        HIP_CHECK(hipMemcpyAsync(af.data(), in_accf, af.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipMemcpyAsync(ag.data(), in_accg, ag.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        const auto vf = inlet_flags_host();
        const double ks = cfg.case_type == CLIP_CASE_DROP ? 1.0 : (cfg.case_type == CLIP_CASE_BUBBLE ? -1.0 : 0.0);
        const double half_length = cfg.reference_length / 2.0;
        double gsig = 0.0;
        for (size_t r = 0; r < inlet_rows.size(); ++r) {
            const long long n = inlet_rows[r] % g.nodes;
            const int qa = (int)(inlet_rows[r] / g.nodes);
            int c[3];
            g.coords(n, c);
            int side = -1;
            for (int s = 2 * L::D - 1; s >= 0; --s) {
                const int a = s >> 1;
                if (cfg.bc_type[s] != CLIP_BC_VELOCITY) continue;
                if (c[a] == ((s & 1) ? g.dmax[a] + 1 : g.dmin[a] - 1)) {
                    if (vf[s][plane_index(g, s, c)]) side = s;
                    break;
                }
            }
            if (side < 0) {
                const double ph0 = (double)adj_phi0[n];
                const double rho = cfg.rho_l + ph0 * (cfg.rho_h - cfg.rho_l);
                gsig += af[r] * ks * ph0 / half_length / (rho / 3.0) * L::w(qa);
                continue;
            }
            const int ax = side >> 1;
            double U[3];
            inlet_velocity(cfg, side, c, U);
            auto dG = [&](int q, double out[3]) {
                double eu = 0.0;
                for (int d = 0; d < L::D; ++d) eu += L::e(q, d) * U[d];
                for (int d = 0; d < 3; ++d)
                    out[d] = d < L::D ? L::w(q) * (3.0 * L::e(q, d) + 9.0 * eu * L::e(q, d) - 3.0 * U[d]) : 0.0;
            };
            const int b = L::opp(qa);
            const double weight = (double)(L::NIN - 1);
            const double phig = (double)adj_phi0[n];
            double da[3], db[3], dfa[3], dga[3], sumt[3] = {0, 0, 0}, sum0[3] = {0, 0, 0};
            dG(qa, da);
            dG(b, db);
            for (int q = 1; q < QL; ++q) {
                if (L::e(q, ax) != 0) continue;
                double dq[3];
                dG(q, dq);
                double et = 0.0;  // e_a,t . e_q,t over the tangential axes
                for (int t = 0; t < L::D; ++t)
                    if (t != ax) et += (double)L::e(qa, t) * L::e(q, t);
                for (int d = 0; d < 3; ++d) {
                    sumt[d] += et * dq[d];
                    sum0[d] += dq[d];
                }
            }
            bool normal = true;
            for (int t = 0; t < L::D; ++t)
                if (t != ax && L::e(qa, t) != 0) normal = false;
            for (int d = 0; d < 3; ++d) {
                dfa[d] = da[d] - db[d] + sumt[d] / weight;
                dga[d] = phig * (da[d] + db[d]) + (normal ? 0.0 : phig * sum0[d] / weight);
            }
            // dU / dA_m = value cos((m+1) theta + phase_m) about geometry 0's centre
            int t1 = (ax + 1) % 3, t2 = (ax + 2) % 3;
            if (L::D == 2) {
                t1 = 1 - ax;
                t2 = -1;
            }
            const double d1 = c[t1] - cfg.geom_center[0][t1];
            const double d2 = (t2 >= 0) ? c[t2] - cfg.geom_center[0][t2] : 0.0;
            const double th = std::atan2(d2, d1);
            for (int m = 0; m < cfg.inlet_modes && m < CLIP_MAX_INLET_MODES; ++m) {
                const double cm = std::cos((m + 1) * th + cfg.inlet_mode_phase[m]);
                double s = 0.0;
                for (int d = 0; d < L::D; ++d) s += (af[r] * dfa[d] + ag[r] * dga[d]) * cfg.bc_value[side][d] * cm;
                adj_inlet_grad[m] += s;
            }
        }
        return gsig;
    }
    void adjoint_inlet_grad(std::vector<double>& out) override {
        if (!adj_grad_ready) throw ConfigError("adjoint: no backward sweep has run");
        out = adj_inlet_grad;
    }

    void uold_soa(std::vector<double>& out) override { download_field(u[1 - ucur], L::D, out); }

    double reduce(int field, unsigned* nonfinite) override {
        HIP_CHECK(hipMemsetAsync(red_bad, 0, sizeof(unsigned), stream));
        const int blocks = 1024;
        k_reduce_interior<CT><<<blocks, 256, 0, stream>>>(phib[phc], g, red_part, red_bad);
        launch_check("k_reduce_interior");
        k_reduce_final<double><<<1, 256, 0, stream>>>(red_part, blocks, red_out);
        launch_check("k_reduce_final");
        double s = 0;
        unsigned bad = 0;
        HIP_CHECK(hipMemcpyAsync(&s, red_out, sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipMemcpyAsync(&bad, red_bad, sizeof(unsigned), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        if (nonfinite) *nonfinite = bad;
        if (field == CLIP_FIELD_RHO) {
            // rho is affine in phi
            const double ni = (double)(g.dmax[0] - g.dmin[0] + 1) * (g.dmax[1] - g.dmin[1] + 1) *
                              (g.dmax[2] - g.dmin[2] + 1);
            return cfg.rho_l * ni + s * (cfg.rho_h - cfg.rho_l);
        }
        return s;
    }
};

}  // namespace clipamd
