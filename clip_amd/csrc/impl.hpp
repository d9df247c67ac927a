// impl.hpp — type-erased solver interface shared by the C-ABI translation
// unit and the per-precision instantiations.
#pragma once

#include "clip_amd.h"
#include "common.hpp"

#include <hip/hip_runtime.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace clipamd {

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
    // deviceInitializer from a given phi (dense SoA, x fastest, ghosts included)
    virtual void init_phi(const std::vector<double>& phi) = 0;
    virtual void step(int n) = 0;
    virtual void get_field(int field, std::vector<double>& out) = 0;  // SoA, ghosts included
    virtual void set_field(int field, const std::vector<double>& in) = 0;
    // Resume from CLIP's file set.  With `phi`/`uold` (the extension file)
    // the pulled buffer, outflow planes, u_old and the phi ghost ring are
    // restored and the step kernels recompute the last macroscopic step
    // exactly; without them the reference's restart is reproduced
    // (f_post read directly, u_old = 0).
    // `raw_f` / `raw_g` (16-bit storage only): the stored bytes of the pulled
    // buffer, used instead of re-encoding f / g, so the restart is exact.
    virtual void resume(const std::vector<double>& f, const std::vector<double>& g, const std::vector<double>& fpost,
                        const std::vector<double>& gpost, const std::vector<double>* fprev,
                        const std::vector<double>* gprev, const std::vector<double>* phi_now,
                        const std::vector<double>* phi_before, const std::vector<double>* uold,
                        const std::vector<unsigned char>* raw_f = nullptr,
                        const std::vector<unsigned char>* raw_g = nullptr) = 0;
    // Restart into the direct state: the reference's arrays as they stand
    // between two solve() calls (f_post / g_post held directly).  `phi_ref`
    // is the phi the raw 16-bit bytes are encoded against (unless ref_zero).
    virtual void resume_direct(const std::vector<double>& f, const std::vector<double>& g,
                               const std::vector<double>& fpost, const std::vector<double>& gpost,
                               const std::vector<double>& phi_now, const std::vector<double>& phi_ref,
                               const std::vector<double>& u, const std::vector<double>& p, bool ref_zero,
                               const std::vector<unsigned char>* raw_f, const std::vector<unsigned char>* raw_g) = 0;
    // phi, the phi the pulled buffer is encoded against, u and p when the
    // handle is in the direct state; false otherwise.
    virtual bool direct_state(std::vector<double>& phi_now, std::vector<double>& phi_ref, std::vector<double>& u,
                              std::vector<double>& p) = 0;
    // NSAllen::macroscopic (NsAllen.cu:866-877) on the current f_post / g_post.
    virtual void macroscopic() = 0;
    // Stored bytes of the pulled (or, in the direct state, the f_post / g_post)
    // buffer of a 16-bit-storage handle; false (and nothing written) otherwise.
    virtual bool raw_pulled(bool is_f, std::vector<unsigned char>& out) = 0;
    // phi of the last macroscopic step and the one before, ghosts included.
    virtual void phi_pair(std::vector<double>& now, std::vector<double>& before) = 0;
    // f_prev / g_prev as CLIP stores them: f_post of the previous step on the
    // outflow planes (the entries the outflow condition reads), the current
    // f_post elsewhere.  SoA, ghosts included.
    virtual void prev_soa(bool is_f, std::vector<double>& out) = 0;
    virtual void uold_soa(std::vector<double>& out) = 0;
    virtual double reduce(int field, unsigned* nonfinite) = 0;
    virtual void sync() = 0;
    virtual void set_stream(hipStream_t s) = 0;
    virtual void* stream_handle() = 0;  // the hipStream_t the handle's work runs on
    virtual void timing(double* total, double* ghost, double* phase, double* fused) const = 0;
    virtual void enable_timing(bool on) = 0;
    virtual void enable_graph(bool on) = 0;
    // Reverse mode from an interior phi0 over nsteps steps; interior arrays
    // are host FP64, x fastest (xslow = 0) or x slowest (xslow = 1).
    virtual void adjoint_forward(const double* phi0, int xslow, int nsteps, int every, double* phi_out) = 0;
    virtual void adjoint_options(double normal_floor) = 0;
    // Inner checkpoint interval of the reverse sweep (0: one level).
    virtual void adjoint_nesting(int inner) = 0;
    // dL/dA_m of the inlet modes after the last adjoint_backward.
    virtual void adjoint_inlet_grad(std::vector<double>& out) = 0;
    virtual void adjoint_backward(const double* grad_out, int xslow, double* grad_phi0, double* grad_sigma,
                                  double* grad_mobility) = 0;
    int D = 2, Q = 9;
    unsigned long long bytes = 0;
    long long step_count = 0;
    bool initialized = false;
    bool direct_old = false;    // pop[1-cur] holds f_post directly
    bool old_ref_zero = false;  // pop[1-cur] stored without phi reference
    std::vector<double> direct_f_interior, direct_g_interior;  // dev_f/dev_g interior after a resume
};

// Factories, one per translation unit (the FP64 one is compiled without
// floating-point contraction so that it reproduces the oracle bit for bit).
std::unique_ptr<Impl> make_impl_f64(const clip_config& c);
std::unique_ptr<Impl> make_impl_f32(const clip_config& c);
std::unique_ptr<Impl> make_impl_mixed(const clip_config& c, bool bf16);

}  // namespace clipamd
