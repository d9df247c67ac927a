// solver.hip — host orchestration of the HIP kernels and the C-ABI of
// include/clip_amd.h.
//
// State of a handle between calls (after clip_init_fresh or clip_step):
//   pop[cur]   f*, g* of the pending collision (the next step pulls it)
//   pop[1-cur] f*, g* the last step pulled (the reference's dev_f / dev_g);
//              right after initialisation or a resume it holds f_post directly
//   u[ucur]    u of the last macroscopic step; u[1-ucur] the one before
//   prev[pcur] outflow planes read by the next pull
//   phi        phi with its ghost ring as the last step's stencils saw it
// so one step is: ghost refresh, k_phase, k_fused (FULL), swap.
#include "clip_amd.h"
#include "common.hpp"
#include "geometry.hpp"
#include "kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <functional>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

namespace clipamd {

namespace {
thread_local std::string g_last_error;
}

int set_error(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

const char* last_error_cstr() { return g_last_error.c_str(); }

struct HipError : std::runtime_error {
    explicit HipError(const std::string& m) : std::runtime_error(m) {}
};

#define HIP_CHECK(expr)                                                                               \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            throw HipError(std::string(#expr) + ": " + hipGetErrorString(e_) + " (" + __FILE__ + ":" + \
                           std::to_string(__LINE__) + ")");                                           \
    } while (0)

// ---------------------------------------------------------------------------
// Type-erased interface over the precision/lattice instantiations.
struct Impl {
    virtual ~Impl() {}
    virtual void init_fresh() = 0;
    virtual void step(int n) = 0;
    virtual void get_field(int field, std::vector<double>& out) = 0;  // SoA, ghosts included
    virtual void set_field(int field, const std::vector<double>& in) = 0;
    virtual void resume(const std::vector<double>& f, const std::vector<double>& g, const std::vector<double>& fpost,
                        const std::vector<double>& gpost, const std::vector<double>* phi,
                        const std::vector<double>* uold) = 0;
    virtual void uold_soa(std::vector<double>& out) = 0;
    virtual double reduce(int field, unsigned* nonfinite) = 0;
    virtual void sync() = 0;
    virtual void set_stream(hipStream_t s) = 0;
    virtual void timing(double* total, double* ghost, double* phase, double* fused) const = 0;
    virtual void enable_timing(bool on) = 0;
    int D = 2, Q = 9;
    unsigned long long bytes = 0;
    long long step_count = 0;
    bool initialized = false;
    bool direct_old = false;  // pop[1-cur] holds f_post directly
    std::vector<double> direct_f_interior, direct_g_interior;  // dev_f/dev_g interior after a resume
};

template <class L, class ST, class CT, class SP>
struct SolverImpl : Impl {
    static constexpr int QL = L::Q;
    clip_config cfg;
    DevGrid g;
    DevBC bc;
    Phys<CT> ph;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    ST* fbuf[2] = {nullptr, nullptr};
    ST* gbuf[2] = {nullptr, nullptr};
    CT* phi = nullptr;
    CT* u[2] = {nullptr, nullptr};
    CT* p = nullptr;
    CT* prevf[2][6] = {};
    CT* prevg[2][6] = {};
    unsigned char* vflag[6] = {};
    CT* scratch = nullptr;        // export staging (max(Q, D) * nodes)
    double* red_part = nullptr;   // reduction partials
    unsigned* red_bad = nullptr;
    double* red_out = nullptr;
    int cur = 1, ucur = 0, pcur = 0;
    bool timing_on = false;
    hipEvent_t ev[8] = {};
    double t_total = 0, t_ghost = 0, t_phase = 0, t_fused = 0;
    std::vector<void*> allocations;

    template <class T>
    T* alloc(size_t count) {
        void* ptr = nullptr;
        if (hipMalloc(&ptr, count * sizeof(T)) != hipSuccess) {
            (void)hipGetLastError();
            throw std::bad_alloc();
        }
        allocations.push_back(ptr);
        bytes += count * sizeof(T);
        return (T*)ptr;
    }

    explicit SolverImpl(const clip_config& c) : cfg(c) {
        D = L::D;
        Q = L::Q;
        if (c.dim != L::D) throw ConfigError("lattice/dimension mismatch");
        // Extents: 2D N+2 with the ghost ring outside N, 3D N with the ghost
        // ring inside N (Domain.cu:22-82).
        for (int a = 0; a < 3; ++a) {
            if (L::D == 2) {
                g.ext[a] = (a < 2) ? (int)c.N[a] + 2 : 1;
                g.dmin[a] = (a < 2) ? 1 : 0;
                g.dmax[a] = (a < 2) ? g.ext[a] - 2 : 0;
            } else {
                g.ext[a] = (int)c.N[a];
                g.dmin[a] = 1;
                g.dmax[a] = g.ext[a] - 2;
            }
        }
        g.sy = g.ext[0];
        g.sz = (long long)g.ext[0] * g.ext[1];
        g.nodes = g.sz * g.ext[2];
        for (int s = 0; s < 6; ++s) bc.type[s] = (s < 2 * L::D) ? c.bc_type[s] : -1;
        bool isWall = false, isFC = false, isNeu = false, isVel = false, isSlip = false;
        for (int s = 0; s < 2 * L::D; ++s) {
            isWall |= c.bc_type[s] == CLIP_BC_WALL;
            isSlip |= c.bc_type[s] == CLIP_BC_SLIP_WALL;
            isFC |= c.bc_type[s] == CLIP_BC_FREE_CONVECT;
            isNeu |= c.bc_type[s] == CLIP_BC_NEUMANN;
            isVel |= c.bc_type[s] == CLIP_BC_VELOCITY;
        }
        for (int a = 0; a < 3; ++a)
            bc.periodic[a] = (a < L::D) && (c.bc_type[2 * a] == CLIP_BC_PERIODIC || c.bc_type[2 * a + 1] == CLIP_BC_PERIODIC);
        bc.is_wall = isWall;
        bc.is_fc = isFC;
        bc.is_neumann = isNeu;
        bc.is_velocity = isVel;
        bc.mirror_gate = isWall || isFC || isSlip || isNeu;
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
            fbuf[b] = alloc<ST>(nodes * QL);
            gbuf[b] = alloc<ST>(nodes * QL);
            u[b] = alloc<CT>(nodes * L::D);
        }
        phi = alloc<CT>(nodes);
        p = alloc<CT>(nodes);
        scratch = alloc<CT>(nodes * std::max(QL, L::D));
        red_part = alloc<double>(1024);
        red_bad = alloc<unsigned>(1);
        red_out = alloc<double>(1);
        for (int s = 0; s < 2 * L::D; ++s) {
            if (c.bc_type[s] == CLIP_BC_FREE_CONVECT)
                for (int b = 0; b < 2; ++b) {
                    prevf[b][s] = alloc<CT>(plane_size(g, s) * QL);
                    prevg[b][s] = alloc<CT>(plane_size(g, s) * QL);
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
        HIP_CHECK(hipMemsetAsync(phi, 0, nodes * sizeof(CT), stream));
        HIP_CHECK(hipMemsetAsync(p, 0, nodes * sizeof(CT), stream));
        for (int b = 0; b < 2; ++b) HIP_CHECK(hipMemsetAsync(u[b], 0, nodes * L::D * sizeof(CT), stream));
        HIP_CHECK(hipStreamSynchronize(stream));
    }

    ~SolverImpl() override {
        if (stream) (void)hipStreamSynchronize(stream);
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

    StepArgs<ST, CT> args(int pull_buf, int uin, int prev_in) {
        StepArgs<ST, CT> a;
        a.g = g;
        a.bc = bc;
        a.ph = ph;
        a.fin = fbuf[pull_buf];
        a.gin = gbuf[pull_buf];
        a.fout = fbuf[1 - pull_buf];
        a.gout = gbuf[1 - pull_buf];
        a.phi = phi;
        a.uin = u[uin];
        a.uout = u[1 - uin];
        a.ucol = u[uin];
        a.p = p;
        for (int s = 0; s < 6; ++s) {
            a.prev_f_in[s] = prevf[prev_in][s];
            a.prev_f_out[s] = prevf[1 - prev_in][s];
            a.prev_g_in[s] = prevg[prev_in][s];
            a.prev_g_out[s] = prevg[1 - prev_in][s];
        }
        return a;
    }

    // Grid over the interior box shrunk by `inset` along each lattice axis.
    dim3 grid_box(int inset) const {
        const int nx = g.dmax[0] - g.dmin[0] + 1 - 2 * inset, ny = g.dmax[1] - g.dmin[1] + 1 - 2 * inset;
        const int nz = g.dmax[2] - g.dmin[2] + 1 - (L::D == 3 ? 2 * inset : 0);
        if (nx <= 0 || ny <= 0 || nz <= 0) return dim3(0, 0, 0);
        return dim3((nx + BX - 1) / BX, (ny + BY - 1) / BY, nz);
    }
    dim3 grid_interior() const { return grid_box(0); }
    unsigned shell_blocks() const { return (unsigned)((shell_count<L>(g) + BX * BY - 1) / (BX * BY)); }

    // Bulk (inner box) and boundary-layer launches of a pulling step kernel.
    template <int MODE>
    void launch_fused_pull(const StepArgs<ST, CT>& a) {
        const dim3 gb = grid_box(1);
        if (gb.x) k_fused<L, ST, CT, SP, false, MODE, false><<<gb, dim3(BX, BY), 0, stream>>>(a);
        k_fused<L, ST, CT, SP, false, MODE, true><<<shell_blocks(), BX * BY, 0, stream>>>(a);
        launch_check("k_fused");
    }
    void launch_phase_pull(const StepArgs<ST, CT>& a) {
        const dim3 gb = grid_box(1);
        if (gb.x) k_phase<L, ST, CT, SP, false, false><<<gb, dim3(BX, BY), 0, stream>>>(a);
        k_phase<L, ST, CT, SP, false, true><<<shell_blocks(), BX * BY, 0, stream>>>(a);
        launch_check("k_phase");
    }

    void launch_check(const char* what) {
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) throw HipError(std::string("launch of ") + what + ": " + hipGetErrorString(e));
    }

    void ghost_refresh() {
        long long total = 0;
        for (int s = 0; s < 2 * L::D; ++s) total += plane_size(g, s);
        const int tb = 256;
        const unsigned blocks = (unsigned)((total + tb - 1) / tb);
        if (!bc.is_velocity) {
            k_ghost_resolve<CT><<<blocks, tb, 0, stream>>>(phi, g, bc, L::D, total, 1, 1);
            launch_check("k_ghost_resolve");
            return;
        }
        k_ghost_resolve<CT><<<blocks, tb, 0, stream>>>(phi, g, bc, L::D, total, 1, 0);
        launch_check("k_ghost_resolve");
        for (int s = 0; s < 2 * L::D; ++s) {
            if (cfg.bc_type[s] != CLIP_BC_VELOCITY) continue;
            const long long ps = plane_size(g, s);
            k_ghost_velocity<CT><<<(unsigned)((ps + tb - 1) / tb), tb, 0, stream>>>(phi, g, vflag[s], s);
            launch_check("k_ghost_velocity");
        }
        if (bc.mirror_gate) {
            k_ghost_resolve<CT><<<blocks, tb, 0, stream>>>(phi, g, bc, L::D, total, 0, 1);
            launch_check("k_ghost_resolve");
        }
    }

    void init_fresh() override {
        // Host phi over every node including ghosts (NsAllen.cu:768-824).
        std::vector<CT> h(g.nodes);
        for (int k = 0; k < g.ext[2]; ++k)
            for (int j = 0; j < g.ext[1]; ++j)
                for (int i = 0; i < g.ext[0]; ++i) h[g.idx(i, j, k)] = (CT)initial_phi(cfg, i, j, k);
        HIP_CHECK(hipMemcpyAsync(phi, h.data(), h.size() * sizeof(CT), hipMemcpyHostToDevice, stream));
        // Distributions, p, u on all nodes of both buffers; outflow planes.
        StepArgs<ST, CT> a = args(1, 1, 1);  // fout = buffer 0, uout = u[0], prev_out = prev[0]
        const int tb = 256;
        k_init<L, ST, CT, SP><<<(unsigned)((g.nodes + tb - 1) / tb), tb, 0, stream>>>(
            a, fbuf[1], gbuf[1], (CT)(cfg.reference_length / 2.0), (CT)cfg.sigma);
        launch_check("k_init");
        // Time-invariant inlet populations inside the disc, side by side.
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
        // First collision straight from the initial f_post = f (buffer 0).
        StepArgs<ST, CT> c = args(0, 0, 1);  // pull 0 -> write 1; u from u[0]; prev_out = prev[0]
        k_fused<L, ST, CT, SP, true, MODE_COLLIDE, false><<<grid_interior(), dim3(BX, BY), 0, stream>>>(c);
        launch_check("k_fused");
        HIP_CHECK(hipStreamSynchronize(stream));
        cur = 1;
        ucur = 0;
        pcur = 0;
        direct_old = true;
        direct_f_interior.clear();
        direct_g_interior.clear();
        step_count = 0;
        initialized = true;
    }

    void step(int n) override {
        t_total = t_ghost = t_phase = t_fused = 0;
        for (int s = 0; s < n; ++s) {
            StepArgs<ST, CT> a = args(cur, ucur, pcur);
            const bool tm = timing_on;
            if (tm) HIP_CHECK(hipEventRecord(ev[0], stream));
            ghost_refresh();
            if (tm) HIP_CHECK(hipEventRecord(ev[1], stream));
            launch_phase_pull(a);
            if (tm) HIP_CHECK(hipEventRecord(ev[2], stream));
            launch_fused_pull<MODE_FULL>(a);
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
            direct_old = false;
            step_count++;
        }
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

    // Exports f_post/g_post (pull of the buffer the last step read, with the
    // u and outflow planes that pull used) into `scratch`; with is_f also
    // refreshes p = sum_q f_post.
    void export_post(bool is_f) {
        const int old = 1 - cur;
        StepArgs<ST, CT> a = args(old, 1 - ucur, 1 - pcur);
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
        switch (field) {
            case CLIP_FIELD_PHI: download(phi, nodes, out); return;
            case CLIP_FIELD_RHO: {
                download(phi, nodes, out);
                for (auto& v : out) v = (double)(ph.rho_l + (CT)v * (ph.rho_h - ph.rho_l));
                return;
            }
            case CLIP_FIELD_U: download(u[ucur], nodes * L::D, out); return;
            case CLIP_FIELD_P: {
                if (!direct_old) {
                    export_post(true);
                }
                download(p, nodes, out);
                return;
            }
            case CLIP_FIELD_MU:
            case CLIP_FIELD_DPHI:
            case CLIP_FIELD_NORMAL: {
                CT* mu = scratch;
                CT* dp = scratch + nodes;
                CT* nr = scratch + nodes * (1 + L::D);
                if (L::Q < 1 + 2 * L::D) throw ConfigError("scratch too small");
                k_export_stencil<L, CT><<<(unsigned)((nodes + tb - 1) / tb), tb, 0, stream>>>(phi, g, ph, mu, dp, nr);
                launch_check("k_export_stencil");
                if (field == CLIP_FIELD_MU) download(mu, nodes, out);
                else if (field == CLIP_FIELD_DPHI) download(dp, nodes * L::D, out);
                else download(nr, nodes * L::D, out);
                return;
            }
            case CLIP_FIELD_F:
            case CLIP_FIELD_G: {
                const int old = 1 - cur;
                const ST* src = (field == CLIP_FIELD_F) ? fbuf[old] : gbuf[old];
                const long long count = (long long)nodes * QL;
                k_export_raw<ST, CT, SP><<<(unsigned)((count + tb - 1) / tb), tb, 0, stream>>>(src, scratch, count, g.nodes);
                launch_check("k_export_raw");
                download(scratch, count, out);
                const auto& keep = (field == CLIP_FIELD_F) ? direct_f_interior : direct_g_interior;
                if (direct_old && !keep.empty()) out = keep;
                if (!direct_old) fill_periodic_ghosts(out, QL);
                return;
            }
            case CLIP_FIELD_F_POST:
            case CLIP_FIELD_G_POST: {
                export_post(field == CLIP_FIELD_F_POST);
                download(scratch, nodes * QL, out);
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
                            a[(size_t)q * g.nodes + g.idx(i, j, k)] = a[(size_t)q * g.nodes + g.idx(s[0], s[1], s[2])];
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

    void upload_pop(ST* dev, const std::vector<double>& in) {
        const long long count = (long long)in.size();
        upload(scratch, in);
        const int tb = 256;
        k_import_raw<ST, CT, SP><<<(unsigned)((count + tb - 1) / tb), tb, 0, stream>>>(dev, scratch, count, g.nodes);
        launch_check("k_import_raw");
        HIP_CHECK(hipStreamSynchronize(stream));
    }

    void set_field(int field, const std::vector<double>& in) override {
        switch (field) {
            case CLIP_FIELD_PHI: upload(phi, in); break;
            case CLIP_FIELD_U: upload(u[ucur], in); break;
            case CLIP_FIELD_P: upload(p, in); break;
            default: throw ConfigError("field cannot be set directly; use a checkpoint resume");
        }
        // Recompute the pending collision from the modified state.
        if (initialized) recollide();
    }

    // Re-runs the collision that produced pop[cur] from the state the last
    // step left behind (used after a field was modified).
    void recollide() {
        const int old = 1 - cur;
        if (!direct_old) export_post(true);  // p from the current f_post
        // pull pop[old] with the u / outflow planes of that pull; collide with u[ucur]
        StepArgs<ST, CT> a = args(old, 1 - ucur, 1 - pcur);
        a.ucol = u[ucur];
        a.uout = nullptr;
        if (direct_old) {
            k_fused<L, ST, CT, SP, true, MODE_COLLIDE, false><<<grid_interior(), dim3(BX, BY), 0, stream>>>(a);
            launch_check("k_fused");
        } else {
            launch_fused_pull<MODE_COLLIDE>(a);
        }
        HIP_CHECK(hipStreamSynchronize(stream));
    }

    void resume(const std::vector<double>& f, const std::vector<double>& gg, const std::vector<double>& fpost,
                const std::vector<double>& gpost, const std::vector<double>* phi_in,
                const std::vector<double>* uold) override {
        // Buffer 0: dev_f / dev_g with the interior replaced by f_post / g_post
        // (read directly by the first kernels); buffer 1: dev_f / dev_g.
        std::vector<double> a0 = f, b0 = gg;
        for (int k = g.dmin[2]; k <= g.dmax[2]; ++k)
            for (int j = g.dmin[1]; j <= g.dmax[1]; ++j)
                for (int i = g.dmin[0]; i <= g.dmax[0]; ++i)
                    for (int q = 0; q < QL; ++q) {
                        const size_t t = (size_t)q * g.nodes + g.idx(i, j, k);
                        a0[t] = fpost[t];
                        b0[t] = gpost[t];
                    }
        upload_pop(fbuf[0], a0);
        upload_pop(gbuf[0], b0);
        upload_pop(fbuf[1], f);
        upload_pop(gbuf[1], gg);
        if (phi_in) upload(phi, *phi_in);
        std::vector<double> zero((size_t)g.nodes * L::D, 0.0);
        upload(u[0], uold ? *uold : zero);
        // macroscopic() of the resume path (main.cu:67-71), then the collision.
        StepArgs<ST, CT> a = args(0, 0, 1);  // pull 0 (direct) -> 1; u_old = u[0] -> u[1]; prev_out = prev[0]
        k_phase<L, ST, CT, SP, true, false><<<grid_interior(), dim3(BX, BY), 0, stream>>>(a);
        launch_check("k_phase");
        if (!phi_in) ghost_refresh();
        k_fused<L, ST, CT, SP, true, MODE_FULL, false><<<grid_interior(), dim3(BX, BY), 0, stream>>>(a);
        launch_check("k_fused");
        HIP_CHECK(hipStreamSynchronize(stream));
        cur = 1;
        ucur = 1;
        pcur = 0;
        direct_old = true;
        direct_f_interior = f;
        direct_g_interior = gg;
        // p of the resumed macroscopic step
        {
            const int tb = 256;
            StepArgs<ST, CT> e = args(0, 0, 1);
            k_export_post<L, ST, CT, SP, true><<<(unsigned)((g.nodes + tb - 1) / tb), tb, 0, stream>>>(e, true, scratch, p);
            launch_check("k_export_post");
            HIP_CHECK(hipStreamSynchronize(stream));
        }
        initialized = true;
    }

    void uold_soa(std::vector<double>& out) override { download(u[1 - ucur], (size_t)g.nodes * L::D, out); }

    double reduce(int field, unsigned* nonfinite) override {
        const CT* src = phi;
        if (field == CLIP_FIELD_RHO) {
            // rho is a function of phi; sum phi and map
            src = phi;
        }
        HIP_CHECK(hipMemsetAsync(red_bad, 0, sizeof(unsigned), stream));
        const int blocks = 1024;
        k_reduce_interior<CT><<<blocks, 256, 0, stream>>>(src, g, L::D, red_part, red_bad, 0.0);
        launch_check("k_reduce_interior");
        k_reduce_final<<<1, 256, 0, stream>>>(red_part, blocks, red_out);
        launch_check("k_reduce_final");
        double s = 0;
        unsigned bad = 0;
        HIP_CHECK(hipMemcpyAsync(&s, red_out, sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipMemcpyAsync(&bad, red_bad, sizeof(unsigned), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        if (nonfinite) *nonfinite = bad;
        if (field == CLIP_FIELD_RHO) {
            const double ni = (double)(g.dmax[0] - g.dmin[0] + 1) * (g.dmax[1] - g.dmin[1] + 1) *
                              (g.dmax[2] - g.dmin[2] + 1);
            return cfg.rho_l * ni + s * (cfg.rho_h - cfg.rho_l);
        }
        return s;
    }
};

std::unique_ptr<Impl> make_impl(const clip_config& c, int precision) {
    if (c.dim == 2) {
        switch (precision) {
            case CLIP_PREC_F64: return std::make_unique<SolverImpl<D2Q9, double, double, StorePlain<double, double>>>(c);
            case CLIP_PREC_F32: return std::make_unique<SolverImpl<D2Q9, float, float, StorePlain<float, float>>>(c);
        }
    } else {
        switch (precision) {
            case CLIP_PREC_F64: return std::make_unique<SolverImpl<D3Q19, double, double, StorePlain<double, double>>>(c);
            case CLIP_PREC_F32: return std::make_unique<SolverImpl<D3Q19, float, float, StorePlain<float, float>>>(c);
        }
    }
    throw ConfigError("unsupported precision " + std::to_string(precision));
}

}  // namespace clipamd

// ===========================================================================
// C-ABI
// ===========================================================================
struct clip_solver {
    clip_config cfg;
    int precision;
    std::unique_ptr<clipamd::Impl> impl;
};

using clipamd::set_error;

namespace {

int guard(const std::function<void()>& fn) {
    try {
        fn();
    } catch (const clipamd::ConfigError& e) {
        return set_error(CLIP_ERR_ARG, e.what());
    } catch (const clipamd::HipError& e) {
        return set_error(CLIP_ERR_HIP, e.what());
    } catch (const std::bad_alloc&) {
        return set_error(CLIP_ERR_NOMEM, "device or host allocation failed");
    } catch (const std::exception& e) {
        return set_error(CLIP_ERR_IO, e.what());
    }
    return CLIP_OK;
}

int ndof_of(const clip_solver* s, int field) {
    const int D = s->cfg.dim, Q = (D == 2) ? 9 : 19;
    switch (field) {
        case CLIP_FIELD_PHI:
        case CLIP_FIELD_RHO:
        case CLIP_FIELD_P:
        case CLIP_FIELD_MU: return 1;
        case CLIP_FIELD_U:
        case CLIP_FIELD_DPHI:
        case CLIP_FIELD_NORMAL: return D;
        case CLIP_FIELD_F:
        case CLIP_FIELD_G:
        case CLIP_FIELD_F_POST:
        case CLIP_FIELD_G_POST: return Q;
    }
    return -1;
}

void extents(const clip_config& c, unsigned ext[3]) {
    for (int a = 0; a < 3; ++a) {
        if (c.dim == 2) ext[a] = (a < 2) ? (unsigned)c.N[a] + 2 : 1;
        else ext[a] = (unsigned)c.N[a];
    }
}

// SoA (dof-major, x fastest) <-> CLIP AoS ((i*ey+j)*ez+k)*ndof+dof.
void soa_to_aos(const std::vector<double>& in, double* out, const unsigned ext[3], int ndof) {
    const size_t nodes = (size_t)ext[0] * ext[1] * ext[2];
    for (unsigned i = 0; i < ext[0]; ++i)
        for (unsigned j = 0; j < ext[1]; ++j)
            for (unsigned k = 0; k < ext[2]; ++k) {
                const size_t aos = ((size_t)i * ext[1] + j) * ext[2] + k;
                const size_t soa = i + (size_t)ext[0] * (j + (size_t)ext[1] * k);
                for (int d = 0; d < ndof; ++d) out[aos * ndof + d] = in[(size_t)d * nodes + soa];
            }
}
void aos_to_soa(const double* in, std::vector<double>& out, const unsigned ext[3], int ndof) {
    const size_t nodes = (size_t)ext[0] * ext[1] * ext[2];
    out.resize(nodes * ndof);
    for (unsigned i = 0; i < ext[0]; ++i)
        for (unsigned j = 0; j < ext[1]; ++j)
            for (unsigned k = 0; k < ext[2]; ++k) {
                const size_t aos = ((size_t)i * ext[1] + j) * ext[2] + k;
                const size_t soa = i + (size_t)ext[0] * (j + (size_t)ext[1] * k);
                for (int d = 0; d < ndof; ++d) out[(size_t)d * nodes + soa] = in[aos * ndof + d];
            }
}

size_t nodes_of(const clip_config& c) {
    unsigned e[3];
    extents(c, e);
    return (size_t)e[0] * e[1] * e[2];
}

}  // namespace

extern "C" {

const char* clip_last_error(void) { return clipamd::last_error_cstr(); }

int clip_create(const clip_config* cfg, int precision, clip_solver** out) {
    if (!cfg || !out) return set_error(CLIP_ERR_ARG, "null argument");
    *out = nullptr;
    clip_config c = *cfg;
    int rc = clip_config_derive(&c);
    if (rc != CLIP_OK) return rc;
    auto s = std::make_unique<clip_solver>();
    s->cfg = c;
    s->precision = precision;
    rc = guard([&] { s->impl = clipamd::make_impl(c, precision); });
    if (rc != CLIP_OK) return rc;
    *out = s.release();
    return CLIP_OK;
}

int clip_set_stream(clip_solver* s, void* stream) {
    if (!s) return set_error(CLIP_ERR_ARG, "null handle");
    return guard([&] { s->impl->set_stream((hipStream_t)stream); });
}

int clip_init_fresh(clip_solver* s) {
    if (!s) return set_error(CLIP_ERR_ARG, "null handle");
    return guard([&] { s->impl->init_fresh(); });
}

int clip_step(clip_solver* s, int nsteps) {
    if (!s) return set_error(CLIP_ERR_ARG, "null handle");
    if (nsteps < 0) return set_error(CLIP_ERR_ARG, "nsteps must be >= 0");
    if (!s->impl->initialized) return set_error(CLIP_ERR_STATE, "clip_step before clip_init_fresh / resume");
    return guard([&] { s->impl->step(nsteps); });
}

int clip_synchronize(clip_solver* s) {
    if (!s) return set_error(CLIP_ERR_ARG, "null handle");
    return guard([&] { s->impl->sync(); });
}

int clip_get_field(clip_solver* s, int field, int layout, double* out) {
    if (!s || !out) return set_error(CLIP_ERR_ARG, "null argument");
    const int nd = ndof_of(s, field);
    if (nd < 0) return set_error(CLIP_ERR_ARG, "unknown field");
    if (layout != CLIP_LAYOUT_AOS && layout != CLIP_LAYOUT_SOA) return set_error(CLIP_ERR_ARG, "unknown layout");
    if (!s->impl->initialized) return set_error(CLIP_ERR_STATE, "solver not initialised");
    return guard([&] {
        std::vector<double> soa;
        s->impl->get_field(field, soa);
        if (layout == CLIP_LAYOUT_SOA) std::memcpy(out, soa.data(), soa.size() * sizeof(double));
        else {
            unsigned e[3];
            extents(s->cfg, e);
            soa_to_aos(soa, out, e, nd);
        }
    });
}

int clip_set_field(clip_solver* s, int field, int layout, const double* in) {
    if (!s || !in) return set_error(CLIP_ERR_ARG, "null argument");
    const int nd = ndof_of(s, field);
    if (nd < 0) return set_error(CLIP_ERR_ARG, "unknown field");
    if (!s->impl->initialized) return set_error(CLIP_ERR_STATE, "solver not initialised");
    return guard([&] {
        std::vector<double> soa;
        unsigned e[3];
        extents(s->cfg, e);
        if (layout == CLIP_LAYOUT_SOA) soa.assign(in, in + nodes_of(s->cfg) * nd);
        else aos_to_soa(in, soa, e, nd);
        s->impl->set_field(field, soa);
    });
}

int clip_sum(clip_solver* s, int field, double* out) {
    if (!s || !out) return set_error(CLIP_ERR_ARG, "null argument");
    if (field != CLIP_FIELD_PHI && field != CLIP_FIELD_RHO) return set_error(CLIP_ERR_ARG, "sum supports PHI and RHO");
    if (!s->impl->initialized) return set_error(CLIP_ERR_STATE, "solver not initialised");
    return guard([&] { *out = s->impl->reduce(field, nullptr); });
}

int clip_check_health(clip_solver* s, double* phi_sum) {
    if (!s) return set_error(CLIP_ERR_ARG, "null handle");
    if (!s->impl->initialized) return set_error(CLIP_ERR_STATE, "solver not initialised");
    unsigned bad = 0;
    double sum = 0;
    int rc = guard([&] { sum = s->impl->reduce(CLIP_FIELD_PHI, &bad); });
    if (rc != CLIP_OK) return rc;
    if (phi_sum) *phi_sum = sum;
    if (bad || !std::isfinite(sum))
        return set_error(CLIP_ERR_DIVERGED, "Solver has diverged: phi holds " + std::to_string(bad) +
                                                " non-finite values");
    return CLIP_OK;
}

int clip_get_info(const clip_solver* s, clip_info* out) {
    if (!s || !out) return set_error(CLIP_ERR_ARG, "null argument");
    std::memset(out, 0, sizeof(*out));
    out->dim = s->cfg.dim;
    out->q = (s->cfg.dim == 2) ? 9 : 19;
    out->precision = s->precision;
    extents(s->cfg, out->ext);
    out->nodes = (unsigned long long)out->ext[0] * out->ext[1] * out->ext[2];
    unsigned long long in = 1;
    for (int a = 0; a < s->cfg.dim; ++a) in *= (unsigned long long)(out->ext[a] - 2);
    out->interior = in;
    out->device_bytes = s->impl->bytes;
    out->step = (unsigned long long)s->impl->step_count;
    return CLIP_OK;
}

int clip_last_step_timing(const clip_solver* s, double* total, double* ghost, double* phase, double* fused) {
    if (!s) return set_error(CLIP_ERR_ARG, "null handle");
    s->impl->timing(total, ghost, phase, fused);
    return CLIP_OK;
}

int clip_enable_timing(clip_solver* s, int on) {
    if (!s) return set_error(CLIP_ERR_ARG, "null handle");
    s->impl->enable_timing(on != 0);
    return CLIP_OK;
}

void clip_destroy(clip_solver* s) { delete s; }

}  // extern "C"

// ---------------------------------------------------------------------------
// Checkpoints: CheckPointer::save / load (src/io/CheckPointer.cu:53-102).
namespace {

template <class T>
void write_raw(const std::string& path, const T* data, size_t count) {
    std::ofstream f(path, std::ios::binary);
    if (!f.is_open()) throw std::runtime_error("Failed to open file for saving: " + path);
    f.write(reinterpret_cast<const char*>(data), (std::streamsize)(count * sizeof(T)));
    if (!f) throw std::runtime_error("write failed: " + path);
}

template <class T>
void read_raw(const std::string& path, T* data, size_t count) {
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) throw std::runtime_error("Failed to open file for reading: " + path);
    f.read(reinterpret_cast<char*>(data), (std::streamsize)(count * sizeof(T)));
    if ((size_t)f.gcount() != count * sizeof(T)) throw std::runtime_error("truncated checkpoint file: " + path);
}

// Population / field file in CLIP_REAL (double for FP64 handles, float otherwise).
void write_field(const std::string& path, const std::vector<double>& aos, bool single) {
    if (!single) return write_raw(path, aos.data(), aos.size());
    std::vector<float> f(aos.begin(), aos.end());
    write_raw(path, f.data(), f.size());
}
void read_field(const std::string& path, std::vector<double>& aos, size_t count, bool single) {
    aos.resize(count);
    if (!single) return read_raw(path, aos.data(), count);
    std::vector<float> f(count);
    read_raw(path, f.data(), count);
    for (size_t i = 0; i < count; ++i) aos[i] = f[i];
}

// TimeInfo::simInfo as laid out on x86-64 (include/utils/TimeInfo.cuh:37-43).
struct SimInfoFile {
    double currentTime;
    int currentStep;
    int pad0;
    double dt;
    unsigned char dtIsSet;
    unsigned char pad1[7];
};
static_assert(sizeof(SimInfoFile) == 32, "simInfo layout");

const char kStateMagic[8] = {'C', 'L', 'I', 'P', 'A', 'M', 'D', '1'};

}  // namespace

extern "C" int clip_checkpoint_rotate(const char* base, unsigned copies) {
    if (!base) return set_error(CLIP_ERR_ARG, "null argument");
    return guard([&] {
        namespace fs = std::filesystem;
        const std::string folder(base);
        if (!fs::exists(folder)) return;
        for (int i = (int)copies - 1; i >= 1; --i) {
            const std::string from = folder + "_" + std::to_string(i), to = folder + "_" + std::to_string(i + 1);
            if (fs::exists(from)) {
                if (fs::exists(to)) fs::remove_all(to);
                fs::rename(from, to);
            }
        }
        const std::string first = folder + "_1";
        if (fs::exists(first)) fs::remove_all(first);
        fs::rename(folder, first);
    });
}

extern "C" int clip_checkpoint_save(clip_solver* s, const char* dir) {
    if (!s || !dir) return set_error(CLIP_ERR_ARG, "null argument");
    if (!s->impl->initialized) return set_error(CLIP_ERR_STATE, "solver not initialised");
    return guard([&] {
        namespace fs = std::filesystem;
        const std::string d(dir);
        fs::create_directories(d);
        unsigned e[3];
        extents(s->cfg, e);
        const size_t nodes = (size_t)e[0] * e[1] * e[2];
        const int Q = (s->cfg.dim == 2) ? 9 : 19, D = s->cfg.dim;
        const bool single = s->precision != CLIP_PREC_F64;
        bool fc = false;
        for (int k = 0; k < 2 * D; ++k) fc |= s->cfg.bc_type[k] == CLIP_BC_FREE_CONVECT;
        {
            std::ofstream info(d + "/info.txt");
            if (!info.is_open()) throw std::runtime_error("Failed to open file for writing: " + d + "/info.txt");
            info << "# Summary Info\n";
            info << "CurrentTime: " << 0 << "\n";
            info << "CurrentStep: " << s->impl->step_count << "\n";
            info << "DomainSize (N): " << s->cfg.N[0] << ", " << s->cfg.N[1] << ", " << s->cfg.N[2] << "\n";
        }
        std::vector<double> soa, aos(nodes * Q);
        const struct {
            int field;
            const char* name;
        } pops[] = {{CLIP_FIELD_G_POST, "dev_g_post"}, {CLIP_FIELD_F_POST, "dev_f_post"}, {CLIP_FIELD_G, "dev_g"},
                    {CLIP_FIELD_F, "dev_f"}};
        for (const auto& pf : pops) {
            s->impl->get_field(pf.field, soa);
            soa_to_aos(soa, aos.data(), e, Q);
            write_field(d + "/" + pf.name + ".bin", aos, single);
        }
        if (fc) {
            // f_prev is overwritten from f_post before its first use after a
            // resume (NsAllen.cu:850-854); the post-streaming arrays are stored.
            s->impl->get_field(CLIP_FIELD_G_POST, soa);
            soa_to_aos(soa, aos.data(), e, Q);
            write_field(d + "/dev_g_prev.bin", aos, single);
            s->impl->get_field(CLIP_FIELD_F_POST, soa);
            soa_to_aos(soa, aos.data(), e, Q);
            write_field(d + "/dev_f_prev.bin", aos, single);
        }
        SimInfoFile ti;
        std::memset(&ti, 0, sizeof(ti));
        ti.currentStep = (int)s->impl->step_count;
        write_raw(d + "/timeInfo.bin", &ti, 1);
        std::vector<double> N(s->cfg.N, s->cfg.N + 3);
        write_field(d + "/domainSize.bin", N, single);
        // Extension (no reference counterpart): phi with its ghost ring and the
        // u that the last macroscopic step started from, for a bit-exact resume.
        std::ofstream st(d + "/clip_amd_state.bin", std::ios::binary);
        if (!st.is_open()) throw std::runtime_error("Failed to open " + d + "/clip_amd_state.bin");
        st.write(kStateMagic, 8);
        const long long hdr[4] = {D, Q, (long long)nodes, s->impl->step_count};
        st.write(reinterpret_cast<const char*>(hdr), sizeof(hdr));
        s->impl->get_field(CLIP_FIELD_PHI, soa);
        st.write(reinterpret_cast<const char*>(soa.data()), (std::streamsize)(soa.size() * sizeof(double)));
        s->impl->uold_soa(soa);
        st.write(reinterpret_cast<const char*>(soa.data()), (std::streamsize)(soa.size() * sizeof(double)));
        if (!st) throw std::runtime_error("write failed: clip_amd_state.bin");
    });
}

extern "C" int clip_checkpoint_load(clip_solver* s, const char* dir) {
    if (!s || !dir) return set_error(CLIP_ERR_ARG, "null argument");
    int rc = guard([&] {
        namespace fs = std::filesystem;
        const std::string d(dir);
        unsigned e[3];
        extents(s->cfg, e);
        const size_t nodes = (size_t)e[0] * e[1] * e[2];
        const int Q = (s->cfg.dim == 2) ? 9 : 19, D = s->cfg.dim;
        const bool single = s->precision != CLIP_PREC_F64;
        std::vector<double> N;
        read_field(d + "/domainSize.bin", N, 3, single);
        for (int a = 0; a < 3; ++a)
            if (N[a] != s->cfg.N[a])
                throw clipamd::ConfigError("Loaded domain size mismatch at dimension " + std::to_string(a) +
                                           ". Expected: " + std::to_string(s->cfg.N[a]) +
                                           ", Found: " + std::to_string(N[a]));
        std::vector<double> aos, fpost, gpost, f, g;
        read_field(d + "/dev_g_post.bin", aos, nodes * Q, single);
        aos_to_soa(aos.data(), gpost, e, Q);
        read_field(d + "/dev_f_post.bin", aos, nodes * Q, single);
        aos_to_soa(aos.data(), fpost, e, Q);
        read_field(d + "/dev_g.bin", aos, nodes * Q, single);
        aos_to_soa(aos.data(), g, e, Q);
        read_field(d + "/dev_f.bin", aos, nodes * Q, single);
        aos_to_soa(aos.data(), f, e, Q);
        SimInfoFile ti;
        read_raw(d + "/timeInfo.bin", &ti, 1);
        std::vector<double> phi, uold;
        bool have_state = false;
        const std::string sp = d + "/clip_amd_state.bin";
        if (fs::exists(sp)) {
            std::ifstream st(sp, std::ios::binary);
            char magic[8];
            long long hdr[4];
            st.read(magic, 8);
            st.read(reinterpret_cast<char*>(hdr), sizeof(hdr));
            if (!st || std::memcmp(magic, kStateMagic, 8) != 0 || hdr[0] != D || hdr[1] != Q ||
                hdr[2] != (long long)nodes)
                throw std::runtime_error("clip_amd_state.bin does not match this configuration");
            phi.resize(nodes);
            uold.resize(nodes * D);
            st.read(reinterpret_cast<char*>(phi.data()), (std::streamsize)(phi.size() * sizeof(double)));
            st.read(reinterpret_cast<char*>(uold.data()), (std::streamsize)(uold.size() * sizeof(double)));
            if (!st) throw std::runtime_error("truncated clip_amd_state.bin");
            have_state = true;
        }
        s->impl->resume(f, g, fpost, gpost, have_state ? &phi : nullptr, have_state ? &uold : nullptr);
        s->impl->step_count = ti.currentStep;
    });
    if (rc == CLIP_ERR_ARG || rc == CLIP_ERR_IO) {
        // domain mismatch and unreadable files are I/O errors of this call
        const std::string msg = clip_last_error();
        return set_error(CLIP_ERR_IO, msg);
    }
    return rc;
}
