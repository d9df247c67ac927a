// solver.hip — the C-ABI of include/clip_amd.h: argument checks, error
// capture, layout conversion (SoA <-> CLIP AoS) and checkpoint files.  The
// device work lives in SolverImpl (solver_impl.hpp), instantiated per
// precision in inst_f64.hip / inst_f32.hip.
#include "clip_amd.h"
#include "common.hpp"
#include "geometry.hpp"
#include "impl.hpp"
#include "launchers.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
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

std::unique_ptr<Impl> make_impl(const clip_config& c, int precision) {
    switch (precision) {
        case CLIP_PREC_F64: return make_impl_f64(c);
        case CLIP_PREC_F32: return make_impl_f32(c);
        case CLIP_PREC_MIXED_F16: return make_impl_mixed(c, false);
        case CLIP_PREC_MIXED_BF16: return make_impl_mixed(c, true);
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

size_t interior_of(const clip_config& c) {
    unsigned e[3];
    extents(c, e);
    size_t n = 1;
    for (int a = 0; a < c.dim; ++a) n *= (size_t)(e[a] - 2);
    return n;
}

}  // namespace

// Solver launchers on caller arrays (CLIP AoS, FP64, device memory).
namespace {

using clipamd::HipError;

bool device_pointer(const void* p) {
    if (!p) return false;
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged;
}

// Runs `body(stream, grid, bc)` for a handle's configuration once the arrays
// are checked to be device memory (optional ones may be null).
int launcher(clip_solver* s, std::initializer_list<const void*> required, std::initializer_list<const void*> optional,
             const std::function<void(hipStream_t, const clipamd::AosGrid&, const clipamd::DevBC&)>& body) {
    if (!s) return set_error(CLIP_ERR_ARG, "null handle");
    for (const void* p : required)
        if (!device_pointer(p)) return set_error(CLIP_ERR_ARG, "launcher arrays must be device memory");
    for (const void* p : optional)
        if (p && !device_pointer(p)) return set_error(CLIP_ERR_ARG, "launcher arrays must be device memory");
    return guard([&] {
        const hipStream_t st = (hipStream_t)s->impl->stream_handle();
        const clipamd::AosGrid g = clipamd::aos_grid(s->cfg);
        const clipamd::DevBC bc = clipamd::bc_of_config(s->cfg, s->cfg.dim);
        body(st, g, bc);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(st));
    });
}

unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

// A device copy of `n` doubles, freed with the object.
struct DevCopy {
    double* p = nullptr;
    DevCopy(const double* src, long long n, hipStream_t st) {
        HIP_CHECK(hipMalloc(&p, (size_t)n * sizeof(double)));
        HIP_CHECK(hipMemcpyAsync(p, src, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, st));
    }
    ~DevCopy() {
        if (p) (void)hipFree(p);
    }
};

template <class Fn>
void with_lattice(int dim, Fn fn) {
    if (dim == 2) fn(clipamd::Lattice<9>{});
    else fn(clipamd::Lattice<19>{});
}

int half_bb(clip_solver* s, bool slip, const double* dev_a, double* dev_a_post, const double* dev_b,
                   double* dev_b_post) {
    if (!dev_b != !dev_b_post) return set_error(CLIP_ERR_ARG, "dev_b and dev_b_post go together");
    return launcher(s, {dev_a, dev_a_post}, {dev_b, dev_b_post},
                    [&](hipStream_t st, const clipamd::AosGrid& g, const clipamd::DevBC& bc) {
                        bool run = false;
                        for (int k = 0; k < 2 * g.D; ++k)
                            run |= s->cfg.bc_type[k] == (slip ? CLIP_BC_SLIP_WALL : CLIP_BC_WALL);
                        if (!run) return;
                        with_lattice(g.D, [&](auto L) {
                            using Lt = decltype(L);
                            clipamd::k_aos_half_bb<<<blocks_for(g.count()), 256, 0, st>>>(
                                dev_a, dev_a_post, dev_b, dev_b_post, g, bc, clipamd::side_lists<Lt>(slip), Lt::Q);
                        });
                    });
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

int clip_init_from_phi(clip_solver* s, const double* phi) {
    if (!s || !phi) return set_error(CLIP_ERR_ARG, "null argument");
    return guard([&] {
        unsigned e[3];
        extents(s->cfg, e);
        std::vector<double> soa;
        aos_to_soa(phi, soa, e, 1);
        s->impl->init_phi(soa);
    });
}

int clip_initial_phi(const clip_config* cfg, double* phi) {
    if (!cfg || !phi) return set_error(CLIP_ERR_ARG, "null argument");
    if (cfg->dim != 2 && cfg->dim != 3) return set_error(CLIP_ERR_ARG, "dim must be 2 or 3");
    if (cfg->num_geometries < 0 || cfg->num_geometries > CLIP_MAX_GEOMETRIES)
        return set_error(CLIP_ERR_ARG, "num_geometries out of range");
    unsigned e[3];
    extents(*cfg, e);
    for (unsigned i = 0; i < e[0]; ++i)
        for (unsigned j = 0; j < e[1]; ++j)
            for (unsigned k = 0; k < e[2]; ++k)
                phi[((size_t)i * e[1] + j) * e[2] + k] = clipamd::initial_phi(*cfg, (int)i, (int)j, (int)k);
    return CLIP_OK;
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
    if (field == CLIP_FIELD_RHO || field == CLIP_FIELD_MU || field == CLIP_FIELD_DPHI || field == CLIP_FIELD_NORMAL)
        return set_error(CLIP_ERR_ARG, "RHO, MU, DPHI and NORMAL are derived from PHI and cannot be set");
    if (layout != CLIP_LAYOUT_AOS && layout != CLIP_LAYOUT_SOA) return set_error(CLIP_ERR_ARG, "unknown layout");
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

int clip_macroscopic(clip_solver* s) {
    if (!s) return set_error(CLIP_ERR_ARG, "null handle");
    if (!s->impl->initialized) return set_error(CLIP_ERR_STATE, "solver not initialised");
    return guard([&] { s->impl->macroscopic(); });
}


int clip_periodic_boundary(clip_solver* s, int ndof, double* dev_a, double* dev_b) {
    if (ndof < 1) return set_error(CLIP_ERR_ARG, "ndof must be positive");
    return launcher(s, {dev_a}, {dev_b}, [&](hipStream_t st, const clipamd::AosGrid& g, const clipamd::DevBC& bc) {
        for (int ax = 0; ax < g.D; ++ax) {
            if (!bc.periodic[ax]) continue;
            const int t1 = (ax == 0) ? 1 : 0, t2 = (ax == 2) ? 1 : 2;
            const long long n = (long long)g.ext[t1] * g.ext[t2] * ndof;
            for (double* arr : {dev_a, dev_b})
                if (arr) clipamd::k_aos_periodic<<<blocks_for(n), 256, 0, st>>>(arr, ndof, g, ax);
        }
    });
}

int clip_wall_boundary(clip_solver* s, const double* dev_a, double* dev_a_post, const double* dev_b,
                       double* dev_b_post) {
    return half_bb(s, false, dev_a, dev_a_post, dev_b, dev_b_post);
}

int clip_slip_wall_boundary(clip_solver* s, const double* dev_a, double* dev_a_post, const double* dev_b,
                            double* dev_b_post) {
    return half_bb(s, true, dev_a, dev_a_post, dev_b, dev_b_post);
}

int clip_free_convect_boundary(clip_solver* s, const double* dev_vel, double* dev_a, const double* dev_a_prev,
                               double* dev_b, const double* dev_b_prev) {
    if (!dev_b != !dev_b_prev) return set_error(CLIP_ERR_ARG, "dev_b and dev_b_prev go together");
    return launcher(s, {dev_vel, dev_a, dev_a_prev}, {dev_b, dev_b_prev},
                    [&](hipStream_t st, const clipamd::AosGrid& g, const clipamd::DevBC& bc) {
                        if (!bc.is_fc) return;
                        with_lattice(g.D, [&](auto L) {
                            using Lt = decltype(L);
                            const long long n = g.count() * Lt::Q;
                            DevCopy ca(dev_a, n, st);
                            std::unique_ptr<DevCopy> cb;
                            if (dev_b) cb.reset(new DevCopy(dev_b, n, st));
                            clipamd::k_aos_free_convect<<<blocks_for(g.count()), 256, 0, st>>>(
                                dev_vel, dev_a, ca.p, dev_a_prev, dev_b, cb ? cb->p : nullptr, dev_b_prev, g, bc,
                                clipamd::side_lists<Lt>(false), Lt::Q);
                            HIP_CHECK(hipStreamSynchronize(st));
                        });
                    });
}

int clip_neumann_boundary(clip_solver* s, double* dev_a, double* dev_b) {
    return launcher(s, {dev_a}, {dev_b}, [&](hipStream_t st, const clipamd::AosGrid& g, const clipamd::DevBC& bc) {
        if (!bc.is_neumann) return;
        with_lattice(g.D, [&](auto L) {
            using Lt = decltype(L);
            const long long n = g.count() * Lt::Q;
            DevCopy ca(dev_a, n, st);
            std::unique_ptr<DevCopy> cb;
            if (dev_b) cb.reset(new DevCopy(dev_b, n, st));
            clipamd::k_aos_neumann<<<blocks_for(g.count()), 256, 0, st>>>(dev_a, ca.p, dev_b, cb ? cb->p : nullptr,
                                                                            g, bc, clipamd::side_lists<Lt>(false),
                                                                            Lt::Q);
            HIP_CHECK(hipStreamSynchronize(st));
        });
    });
}

int clip_velocity_boundary(clip_solver* s, double* dev_c, double* dev_f, double* dev_g) {
    return launcher(s, {dev_c, dev_f, dev_g}, {}, [&](hipStream_t st, const clipamd::AosGrid& g,
                                                     const clipamd::DevBC& bc) {
        if (!bc.is_velocity) return;
        with_lattice(g.D, [&](auto L) {
            using Lt = decltype(L);
            for (int side = 0; side < 2 * g.D; ++side) {
                if (s->cfg.bc_type[side] != CLIP_BC_VELOCITY) continue;
                // disc flags and inlet velocity of every ghost-plane cell (t1 fastest)
                const int ax = side / 2, t1 = (ax == 0) ? 1 : 0, t2 = (ax == 2) ? 1 : 2;
                const long long cells = (long long)g.ext[t1] * g.ext[t2];
                std::vector<unsigned char> flag(cells);
                std::vector<double> U(3 * cells);
                for (long long t = 0; t < cells; ++t) {
                    int cc[3];
                    cc[t1] = (int)(t % g.ext[t1]);
                    cc[t2] = (int)(t / g.ext[t1]);
                    cc[ax] = (side & 1) ? g.dmax[ax] + 1 : g.dmin[ax] - 1;
                    const double z = (g.D == 3) ? (double)cc[2] : 0.0;
                    flag[t] = clipamd::sdf(s->cfg, 0, cc[0], cc[1], z) <= 0 ? 1 : 0;
                    clipamd::inlet_velocity(s->cfg, side, cc, &U[3 * t]);
                }
                unsigned char* dflag = nullptr;
                double* dU = nullptr;
                HIP_CHECK(hipMalloc(&dflag, flag.size()));
                if (hipMalloc(&dU, U.size() * sizeof(double)) != hipSuccess) {
                    (void)hipFree(dflag);
                    throw clipamd::HipError("hipMalloc of the inlet velocity plane failed");
                }
                hipError_t e = hipMemcpyAsync(dflag, flag.data(), flag.size(), hipMemcpyHostToDevice, st);
                if (e == hipSuccess)
                    e = hipMemcpyAsync(dU, U.data(), U.size() * sizeof(double), hipMemcpyHostToDevice, st);
                if (e == hipSuccess) {
                    clipamd::k_aos_velocity<Lt><<<blocks_for(cells), 256, 0, st>>>(
                        dev_c, dev_f, dev_g, g, side, dflag, dU, clipamd::side_lists<Lt>(false));
                    e = hipStreamSynchronize(st);
                }
                (void)hipFree(dflag);
                (void)hipFree(dU);
                HIP_CHECK(e);
            }
        });
    });
}

int clip_mirror_boundary(clip_solver* s, double* dev_a) {
    return launcher(s, {dev_a}, {}, [&](hipStream_t st, const clipamd::AosGrid& g, const clipamd::DevBC& bc) {
        if (!bc.mirror_gate) return;
        for (int side = 0; side < 2 * g.D; ++side) {
            const int t = s->cfg.bc_type[side];
            if (!(t == CLIP_BC_WALL || t == CLIP_BC_FREE_CONVECT || t == CLIP_BC_NEUMANN || t == CLIP_BC_SLIP_WALL ||
                  t == CLIP_BC_DO_NOTHING))
                continue;
            const int ax = side / 2, t1 = (ax == 0) ? 1 : 0, t2 = (ax == 2) ? 1 : 2;
            clipamd::k_aos_mirror<<<blocks_for((long long)g.ext[t1] * g.ext[t2]), 256, 0, st>>>(dev_a, g, side);
        }
    });
}

int clip_copy_rate(unsigned long long bytes, int reps, double* gbs) {
    if (!gbs || reps < 1 || bytes < 16) return set_error(CLIP_ERR_ARG, "bad argument");
    return guard([&] {
        const size_t n = (size_t)(bytes / 16);
        char* buf = nullptr;
        HIP_CHECK(hipMalloc(&buf, 2 * n * 16));
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hipError_t e = hipEventCreate(&e0);
        if (e == hipSuccess) e = hipEventCreate(&e1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipMemset(buf, 1, 2 * n * 16);
        const unsigned blocks = (unsigned)((n + 255) / 256);
        const auto* src = reinterpret_cast<const clipamd::Copy16*>(buf);
        auto* dst = reinterpret_cast<clipamd::Copy16*>(buf + n * 16);
        if (e == hipSuccess) {
            clipamd::k_copy16<<<blocks, 256>>>(src, dst, n);  // warm-up
            e = hipEventRecord(e0);
        }
        for (int r = 0; r < reps && e == hipSuccess; ++r) clipamd::k_copy16<<<blocks, 256>>>(src, dst, n);
        if (e == hipSuccess) e = hipEventRecord(e1);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        (void)hipFree(buf);
        HIP_CHECK(e);
        *gbs = 2.0 * (double)n * 16.0 / ((double)ms / reps * 1e-3) / 1e9;
    });
}

int clip_geometry_sdf(const clip_config* cfg, int id, double x, double y, double z, double* out) {
    if (!cfg || !out) return set_error(CLIP_ERR_ARG, "null argument");
    if (cfg->num_geometries < 0 || cfg->num_geometries > CLIP_MAX_GEOMETRIES)
        return set_error(CLIP_ERR_ARG, "num_geometries out of range");
    *out = clipamd::sdf(*cfg, id, x, y, z);
    return CLIP_OK;
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

int clip_enable_graph(clip_solver* s, int on) {
    if (!s) return set_error(CLIP_ERR_ARG, "null handle");
    s->impl->enable_graph(on != 0);
    return CLIP_OK;
}

int clip_enable_timing(clip_solver* s, int on) {
    if (!s) return set_error(CLIP_ERR_ARG, "null handle");
    s->impl->enable_timing(on != 0);
    return CLIP_OK;
}

int clip_adjoint_forward(clip_solver* s, const double* phi0, int nsteps, int checkpoint_every, double* phi_out,
                         int layout) {
    if (!s || !phi0 || !phi_out) return set_error(CLIP_ERR_ARG, "null argument");
    if (nsteps < 1 || checkpoint_every < 1) return set_error(CLIP_ERR_ARG, "nsteps and checkpoint_every must be >= 1");
    if (layout != CLIP_INTERIOR_X_FASTEST && layout != CLIP_INTERIOR_X_SLOWEST)
        return set_error(CLIP_ERR_ARG, "unknown interior layout");
    return guard([&] { s->impl->adjoint_forward(phi0, layout, nsteps, checkpoint_every, phi_out); });
}

int clip_adjoint_backward(clip_solver* s, const double* grad_phi_out, double* grad_phi0, double* grad_sigma,
                          double* grad_mobility, int layout) {
    if (!s || !grad_phi_out || !grad_phi0) return set_error(CLIP_ERR_ARG, "null argument");
    if (layout != CLIP_INTERIOR_X_FASTEST && layout != CLIP_INTERIOR_X_SLOWEST)
        return set_error(CLIP_ERR_ARG, "unknown interior layout");
    return guard([&] { s->impl->adjoint_backward(grad_phi_out, layout, grad_phi0, grad_sigma, grad_mobility); });
}

int clip_adjoint_set_nesting(clip_solver* s, int inner_every) {
    if (!s) return set_error(CLIP_ERR_ARG, "null handle");
    return guard([&] { s->impl->adjoint_nesting(inner_every); });
}

int clip_adjoint_inlet_gradient(clip_solver* s, double* grad, int n) {
    if (!s || !grad) return set_error(CLIP_ERR_ARG, "null argument");
    if (n < 0) return set_error(CLIP_ERR_ARG, "n must be >= 0");
    int rc = guard([&] {
        std::vector<double> gr;
        s->impl->adjoint_inlet_grad(gr);
        for (int m = 0; m < n; ++m) grad[m] = m < (int)gr.size() ? gr[m] : 0.0;
    });
    if (rc == CLIP_ERR_ARG) return set_error(CLIP_ERR_STATE, clip_last_error());
    return rc;
}

int clip_adjoint_set_normal_floor(clip_solver* s, double normal_floor) {
    if (!s) return set_error(CLIP_ERR_ARG, "null handle");
    return guard([&] { s->impl->adjoint_options(normal_floor); });
}

// VTSwriter::writeVTSBinaryFile (VTSwriter.cu:84-205): one ASCII VTK
// StructuredGrid file over the interior nodes with Velocity, C, P and Rho.
// The fields leave the device only here (the reference copies them to the
// host every step, VTSwriter.cu:72-79).
int clip_write_vts(clip_solver* s, const char* path) {
    if (!s || !path) return set_error(CLIP_ERR_ARG, "null argument");
    if (!s->impl->initialized) return set_error(CLIP_ERR_STATE, "solver not initialised");
    return guard([&] {
        const clip_config& c = s->cfg;
        const int D = c.dim;
        unsigned e[3];
        extents(c, e);
        std::vector<double> phi, p, rho, u;
        s->impl->get_field(CLIP_FIELD_PHI, phi);
        s->impl->get_field(CLIP_FIELD_P, p);
        s->impl->get_field(CLIP_FIELD_RHO, rho);
        s->impl->get_field(CLIP_FIELD_U, u);
        const size_t nodes = (size_t)e[0] * e[1] * e[2];
        unsigned lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = (a < D) ? 1 : 0;
            hi[a] = (a < D) ? e[a] - 2 : 0;
        }
        std::ofstream f(path);
        if (!f) throw std::runtime_error(std::string("cannot open ") + path);
        char ext[160];
        std::snprintf(ext, sizeof(ext), "%u %u %u %u %u %u", lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]);
        f << "<?xml version=\"1.0\"?>\n";
        f << "<VTKFile type=\"StructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
        f << "<StructuredGrid WholeExtent=\"" << ext << "\">\n";
        f << "<Piece Extent=\"" << ext << "\">\n";
        auto each = [&](const std::function<void(size_t)>& fn) {
            for (unsigned k = lo[2]; k <= hi[2]; ++k)
                for (unsigned j = lo[1]; j <= hi[1]; ++j)
                    for (unsigned i = lo[0]; i <= hi[0]; ++i) fn(i + (size_t)e[0] * (j + (size_t)e[1] * k));
        };
        f << "<Points>\n<DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n";
        for (unsigned k = lo[2]; k <= hi[2]; ++k)
            for (unsigned j = lo[1]; j <= hi[1]; ++j)
                for (unsigned i = lo[0]; i <= hi[0]; ++i) f << double(i) << " " << double(j) << " " << double(k) << "\n";
        f << "</DataArray>\n</Points>\n";
        f << "<PointData Scalars=\"scalars\" Vectors=\"velocity\">\n";
        f << "<DataArray type=\"Float64\" Name=\"Velocity\" NumberOfComponents=\"" << D << "\" format=\"ascii\">\n";
        each([&](size_t n) {
            f << u[n] << " " << u[nodes + n];
            if (D == 3) f << " " << u[2 * nodes + n];
            f << "\n";
        });
        f << "</DataArray>\n";
        const std::pair<const char*, const std::vector<double>*> scal[3] = {{"C", &phi}, {"P", &p}, {"Rho", &rho}};
        for (const auto& sc : scal) {
            f << "<DataArray type=\"Float64\" Name=\"" << sc.first << "\" format=\"ascii\">\n";
            each([&](size_t n) { f << (*sc.second)[n] << "\n"; });
            f << "</DataArray>\n";
        }
        f << "</PointData>\n</Piece>\n</StructuredGrid>\n</VTKFile>\n";
        if (!f) throw std::runtime_error(std::string("write failed: ") + path);
    });
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

// version 2: pulled state only; version 3 adds the mode word (0 pulled, 1 direct)
const char kStateMagic2[8] = {'C', 'L', 'I', 'P', 'A', 'M', 'D', '2'};
const char kStateMagic[8] = {'C', 'L', 'I', 'P', 'A', 'M', 'D', '3'};
// optional trailing block of clip_amd_state.bin: raw 16-bit pulled populations
const char kRawTag[8] = {'R', 'A', 'W', 'P', 'O', 'P', '1', '6'};

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
            // f_post of the previous step on the outflow planes (what the
            // outflow condition reads), the current f_post elsewhere.
            s->impl->prev_soa(false, soa);
            soa_to_aos(soa, aos.data(), e, Q);
            write_field(d + "/dev_g_prev.bin", aos, single);
            s->impl->prev_soa(true, soa);
            soa_to_aos(soa, aos.data(), e, Q);
            write_field(d + "/dev_f_prev.bin", aos, single);
        }
        SimInfoFile ti;
        std::memset(&ti, 0, sizeof(ti));
        ti.currentStep = (int)s->impl->step_count;
        write_raw(d + "/timeInfo.bin", &ti, 1);
        std::vector<double> N(s->cfg.N, s->cfg.N + 3);
        write_field(d + "/domainSize.bin", N, single);
        // Extension (no reference counterpart).  Pulled state (mode 0): phi of
        // the last and the previous macroscopic step with their ghost rings,
        // and the u the last macroscopic step started from, for a bit-exact
        // resume.  Direct state (mode 1, right after initialisation, a field
        // upload or macroscopic()): phi, the phi the stored f_post / g_post
        // are encoded against, u, p and whether that encoding is absolute.
        std::ofstream st(d + "/clip_amd_state.bin", std::ios::binary);
        if (!st.is_open()) throw std::runtime_error("Failed to open " + d + "/clip_amd_state.bin");
        std::vector<double> now, ref, uu, pp;
        const bool direct = s->impl->direct_state(now, ref, uu, pp);
        st.write(kStateMagic, 8);
        const long long hdr[5] = {D, Q, (long long)nodes, s->impl->step_count, direct ? 1 : 0};
        st.write(reinterpret_cast<const char*>(hdr), sizeof(hdr));
        auto put = [&](const std::vector<double>& v) {
            st.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
        };
        if (direct) {
            put(now);
            put(ref);
            put(uu);
            put(pp);
            const long long zero = s->impl->old_ref_zero ? 1 : 0;
            st.write(reinterpret_cast<const char*>(&zero), sizeof(zero));
        } else {
            std::vector<double> before;
            s->impl->phi_pair(soa, before);
            put(soa);
            put(before);
            s->impl->uold_soa(soa);
            put(soa);
        }
        // 16-bit storage: the stored bytes of the pulled f and g buffers, so
        // that the restart does not re-encode decoded values
        std::vector<unsigned char> rf, rg;
        if (s->impl->raw_pulled(true, rf) && s->impl->raw_pulled(false, rg)) {
            st.write(kRawTag, 8);
            const long long head[2] = {(long long)s->precision, (long long)rf.size()};  // storage type, bytes
            st.write(reinterpret_cast<const char*>(head), sizeof(head));
            st.write(reinterpret_cast<const char*>(rf.data()), (std::streamsize)rf.size());
            st.write(reinterpret_cast<const char*>(rg.data()), (std::streamsize)rg.size());
        }
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
        std::vector<double> fprev, gprev;
        bool have_prev = false;
        if (fs::exists(d + "/dev_f_prev.bin") && fs::exists(d + "/dev_g_prev.bin")) {
            read_field(d + "/dev_f_prev.bin", aos, nodes * Q, single);
            aos_to_soa(aos.data(), fprev, e, Q);
            read_field(d + "/dev_g_prev.bin", aos, nodes * Q, single);
            aos_to_soa(aos.data(), gprev, e, Q);
            have_prev = true;
        }
        SimInfoFile ti;
        read_raw(d + "/timeInfo.bin", &ti, 1);
        std::vector<double> phi, phi_before, uold, pp;
        std::vector<unsigned char> raw_f, raw_g;
        bool have_state = false, have_raw = false, direct = false, ref_zero = true;
        const std::string sp = d + "/clip_amd_state.bin";
        if (fs::exists(sp)) {
            std::ifstream st(sp, std::ios::binary);
            char magic[8];
            long long hdr[5] = {0, 0, 0, 0, 0};
            st.read(magic, 8);
            const bool v3 = st && std::memcmp(magic, kStateMagic, 8) == 0;
            const bool v2 = st && std::memcmp(magic, kStateMagic2, 8) == 0;
            st.read(reinterpret_cast<char*>(hdr), (std::streamsize)((v3 ? 5 : 4) * sizeof(long long)));
            if (!st || !(v2 || v3) || hdr[0] != D || hdr[1] != Q || hdr[2] != (long long)nodes || hdr[4] < 0 ||
                hdr[4] > 1)
                throw std::runtime_error("clip_amd_state.bin does not match this configuration");
            direct = hdr[4] == 1;
            auto get = [&](std::vector<double>& v, size_t n) {
                v.resize(n);
                st.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(double)));
            };
            get(phi, nodes);
            get(phi_before, nodes);
            get(uold, nodes * D);
            if (direct) {
                get(pp, nodes);
                long long zero = 1;
                st.read(reinterpret_cast<char*>(&zero), sizeof(zero));
                ref_zero = zero != 0;
            }
            if (!st) throw std::runtime_error("truncated clip_amd_state.bin");
            have_state = true;
            char tag[8];
            if (s->precision == CLIP_PREC_MIXED_F16 || s->precision == CLIP_PREC_MIXED_BF16) {
                st.read(tag, 8);
                if (st && std::memcmp(tag, kRawTag, 8) == 0) {
                    long long head[2] = {-1, 0};  // storage type, bytes
                    st.read(reinterpret_cast<char*>(head), sizeof(head));
                    if (!st || head[1] != (long long)(nodes * Q * 2))
                        throw std::runtime_error("clip_amd_state.bin: raw population block does not match");
                    raw_f.resize((size_t)head[1]);
                    raw_g.resize((size_t)head[1]);
                    st.read(reinterpret_cast<char*>(raw_f.data()), head[1]);
                    st.read(reinterpret_cast<char*>(raw_g.data()), head[1]);
                    if (!st) throw std::runtime_error("truncated clip_amd_state.bin");
                    // bytes of the other 16-bit type are not these populations:
                    // then the decoded FP64 files are re-encoded instead
                    have_raw = head[0] == (long long)s->precision;
                }
            }
        }
        if (have_state && direct)
            s->impl->resume_direct(f, g, fpost, gpost, phi, phi_before, uold, pp, ref_zero,
                                   have_raw ? &raw_f : nullptr, have_raw ? &raw_g : nullptr);
        else
            s->impl->resume(f, g, fpost, gpost, have_prev ? &fprev : nullptr, have_prev ? &gprev : nullptr,
                            have_state ? &phi : nullptr, have_state ? &phi_before : nullptr,
                            have_state ? &uold : nullptr, have_raw ? &raw_f : nullptr, have_raw ? &raw_g : nullptr);
        s->impl->step_count = ti.currentStep;
    });
    if (rc == CLIP_ERR_ARG || rc == CLIP_ERR_IO) {
        // domain mismatch and unreadable files are I/O errors of this call
        const std::string msg = clip_last_error();
        return set_error(CLIP_ERR_IO, msg);
    }
    return rc;
}
