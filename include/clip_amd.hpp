// clip_amd.hpp — C++ host surface over the C-ABI, with the reference's class
// names so that code written against CLIP's host classes maps one to one:
//
//   clip::InputData     include/input/InputData.cuh:42-151   (config.txt -> SimParams)
//   clip::Domain        include/domain/Domain.cuh:30-112      (extents, ghost ring)
//   clip::Boundary      include/boundary/Boundary.cuh:48-190  (BC map and flags)
//   clip::Geometry      include/geometry/Geometry.cuh         (geometry pool)
//   clip::DataArray     include/data/DataArray.cuh            (device / host arrays)
//   clip::Solver        include/solver/Solver.cuh:37-151      (boundary launchers)
//   clip::NSAllen       include/equation/NsAllen.cuh:50-95    (initialCondition,
//                       deviceInitializer, solve, macroscopic)
//   clip::TimeInfo      include/utils/TimeInfo.cuh:30-97
//   clip::Reporter      include/utils/Reporter.cuh            (Sum(c), divergence)
//   clip::CheckPointer  include/io/CheckPointer.cuh           (save / load / rotation)
//
// Errors throw std::runtime_error carrying clip_last_error() (the reference's
// Logger::Error also throws).  All device work runs in libclip_amd.so.
#pragma once

#include "clip_amd.h"

#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <filesystem>
#include <stdexcept>
#include <string>
#include <vector>

namespace clip {

inline void check(int rc) {
    if (rc != CLIP_OK) throw std::runtime_error(std::string("clip: ") + clip_last_error());
}

class InputData {
public:
    explicit InputData(const std::string& filename = "config.txt", int dim_override = 0) : m_filename(filename) {
        check(clip_config_from_file(filename.c_str(), dim_override, &params));
    }
    std::string getConfig() const { return m_filename; }
    clip_config params;

private:
    std::string m_filename;
};

class Domain {
public:
    explicit Domain(const InputData& idata) {
        const clip_config& c = idata.params;
        for (int a = 0; a < 3; ++a) {
            if (c.dim == 2) {
                extent[a] = (a < 2) ? (unsigned)c.N[a] + 2 : 1;
                domainMinIdx[a] = (a < 2) ? 1 : 0;
                domainMaxIdx[a] = (a < 2) ? extent[a] - 2 : 0;
            } else {
                extent[a] = (unsigned)c.N[a];
                domainMinIdx[a] = 1;
                domainMaxIdx[a] = extent[a] - 2;
            }
            ghostDomainMinIdx[a] = 0;
            ghostDomainMaxIdx[a] = extent[a] - 1;
        }
        domainSize = (size_t)extent[0] * extent[1] * extent[2];
    }
    // CLIP's AoS flat index (Domain.cuh:63-68).
    size_t getIndex(unsigned i, unsigned j, unsigned k, unsigned ndof = 1, unsigned dof = 0) const {
        return (((size_t)i * extent[1] + j) * extent[2] + k) * ndof + dof;
    }
    unsigned extent[3], domainMinIdx[3], domainMaxIdx[3], ghostDomainMinIdx[3], ghostDomainMaxIdx[3];
    size_t domainSize;
};

class Boundary {
public:
    Boundary(const InputData& idata, const Domain&) {
        const clip_config& c = idata.params;
        for (int s = 0; s < 2 * c.dim; ++s) {
            types[s] = c.bc_type[s];
            isPeriodic |= types[s] == CLIP_BC_PERIODIC;
            isWall |= types[s] == CLIP_BC_WALL;
            isSlipWall |= types[s] == CLIP_BC_SLIP_WALL;
            isFreeConvect |= types[s] == CLIP_BC_FREE_CONVECT;
            isNeumann |= types[s] == CLIP_BC_NEUMANN;
            isVelocity |= types[s] == CLIP_BC_VELOCITY;
        }
    }
    int types[6] = {CLIP_BC_UNKNOWN, CLIP_BC_UNKNOWN, CLIP_BC_UNKNOWN, CLIP_BC_UNKNOWN, CLIP_BC_UNKNOWN, CLIP_BC_UNKNOWN};
    bool isPeriodic = false, isWall = false, isSlipWall = false, isFreeConvect = false, isNeumann = false,
         isVelocity = false;
};

class Geometry {
public:
    explicit Geometry(const InputData& idata) : numGeometries(idata.params.num_geometries), m_params(idata.params) {}
    // Geometry::sdf (Geometry.cuh:102-161): signed distance to the first
    // object carrying `id`, positive outside (1e10 when no object has it).
    double sdf(int id, double x, double y, double z = 0.0) const {
        double d = 0.0;
        check(clip_geometry_sdf(&m_params, id, x, y, z, &d));
        return d;
    }
    int numGeometries;

private:
    clip_config m_params;
};

class TimeInfo {
public:
    explicit TimeInfo(const InputData& idata) : m_final(idata.params.final_step) {}
    int getCurrentStep() const { return m_step; }
    double getFinalStep() const { return m_final; }
    void increment() { ++m_step; }
    void setCurrentStep(int s) { m_step = s; }

private:
    int m_step = 0;
    unsigned m_final;
};

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("clip: ") + what + ": " + hipGetErrorString(e));
}

class NSAllen;

// DataArray (include/data/DataArray.cuh, src/data/DataArray.cu:48-138): the
// reference's device and host arrays, FP64 in CLIP's AoS layout with the ghost
// ring.  The solver keeps its own state in its device layout; an NSAllen built
// on a DataArray fills host_c in initialCondition(), starts from dev_c in
// deviceInitializer() and writes its fields into these arrays on
// updateHost() and syncFromSolver().  The boundary launchers of clip::Solver
// work on them (or on any arrays of this layout).
class DataArray {
public:
    DataArray(const InputData& idata, const Domain& domain, const Boundary& boundary)
        : m_params(idata.params), m_nodes(domain.domainSize), m_fc(boundary.isFreeConvect) {}
    ~DataArray() { release(); }
    DataArray(const DataArray&) = delete;
    DataArray& operator=(const DataArray&) = delete;

    struct deviceDataArray {
        double *dev_f = nullptr, *dev_f_post = nullptr, *dev_g = nullptr, *dev_g_post = nullptr;
        double *dev_rho = nullptr, *dev_p = nullptr, *dev_c = nullptr, *dev_dc = nullptr, *dev_vel = nullptr;
        double *dev_mu = nullptr, *dev_normal = nullptr, *dev_f_prev = nullptr, *dev_g_prev = nullptr;
    } deviceDA;
    struct hostDataArray {
        double *host_c = nullptr, *host_p = nullptr, *host_vel = nullptr, *host_rho = nullptr;
        double *host_f = nullptr, *host_g = nullptr, *host_f_post = nullptr, *host_g_post = nullptr;
        double *host_f_prev = nullptr, *host_g_prev = nullptr;
    } hostDA;

    // DataArray::createVectors (DataArray.cu:48-121).
    void createVectors() {
        release();
        const int Q = q(), D = m_params.dim;
        if (m_fc) {
            dev(deviceDA.dev_f_prev, Q);
            dev(deviceDA.dev_g_prev, Q);
        }
        dev(deviceDA.dev_f, Q);
        dev(deviceDA.dev_g, Q);
        dev(deviceDA.dev_f_post, Q);
        dev(deviceDA.dev_g_post, Q);
        dev(deviceDA.dev_rho, 1);
        dev(deviceDA.dev_mu, 1);
        dev(deviceDA.dev_c, 1);
        dev(deviceDA.dev_p, 1);
        dev(deviceDA.dev_vel, D);
        dev(deviceDA.dev_dc, D);
        dev(deviceDA.dev_normal, D);
        host(hostDA.host_c, 1);
        host(hostDA.host_p, 1);
        host(hostDA.host_vel, D);
        host(hostDA.host_rho, 1);
        host(hostDA.host_f, Q);
        host(hostDA.host_g, Q);
        host(hostDA.host_f_post, Q);
        host(hostDA.host_g_post, Q);
        if (m_fc) {
            host(hostDA.host_f_prev, Q);
            host(hostDA.host_g_prev, Q);
        }
    }
    // DataArray::updateDevice (DataArray.cu:123-127): host_c -> dev_c.
    void updateDevice() {
        if (!hostDA.host_c) throw std::runtime_error("clip: DataArray::createVectors was not called");
        hip_check(hipMemcpy(deviceDA.dev_c, hostDA.host_c, m_nodes * sizeof(double), hipMemcpyHostToDevice),
                  "DataArray::updateDevice");
    }
    // DataArray::updateHost (DataArray.cu:129-138): dev_c, dev_p, dev_rho,
    // dev_vel -> host, after the bound solver wrote its fields into them.
    inline void updateHost();
    // The bound solver's fields into the device arrays dev_f, dev_g,
    // dev_f_post, dev_g_post, dev_c, dev_rho, dev_p, dev_vel, dev_mu, dev_dc
    // and dev_normal (and the populations into the host arrays).  dev_f_prev
    // / dev_g_prev are the caller's: the solver keeps only the outflow
    // planes of them, which its checkpoint files carry.
    inline void syncFromSolver();

    size_t domainSize() const { return m_nodes; }
    int q() const { return m_params.dim == 2 ? 9 : 19; }
    NSAllen* solver() const { return m_solver; }
    void bind(NSAllen* ns) { m_solver = ns; }

private:
    void dev(double*& p, int ndof) {
        hip_check(hipMalloc((void**)&p, m_nodes * ndof * sizeof(double)), "hipMalloc (DataArray)");
        m_dev.push_back(p);
    }
    void host(double*& p, int ndof) {
        m_host.emplace_back(m_nodes * ndof, 0.0);
        p = m_host.back().data();
    }
    void release() {
        for (double* p : m_dev) (void)hipFree(p);
        m_dev.clear();
        m_host.clear();
        deviceDA = deviceDataArray{};
        hostDA = hostDataArray{};
    }
    clip_config m_params;
    size_t m_nodes;
    bool m_fc;
    NSAllen* m_solver = nullptr;
    std::vector<double*> m_dev;
    std::vector<std::vector<double>> m_host;
};

// Solver (include/solver/Solver.cuh:37-151): the boundary launchers on device
// arrays in CLIP's layout, for this solver's configuration.  The template
// arguments are the reference's (Q = values per node, dof = in-directions per
// side) and are checked against the configuration.
class Solver {
public:
    Solver(const InputData& idata, const Domain& domain, const Boundary& boundary, const Geometry& geom,
           int precision = CLIP_PREC_F64)
        : m_idata(&idata), m_domain(&domain), m_boundary(&boundary), m_geom(&geom) {
        check(clip_create(&idata.params, precision, &m_s));
    }
    virtual ~Solver() { clip_destroy(m_s); }
    Solver(const Solver&) = delete;
    Solver& operator=(const Solver&) = delete;

    template <unsigned Q>
    void periodicBoundary(double* dev_a, double* dev_b = nullptr) {
        check(clip_periodic_boundary(m_s, (int)Q, dev_a, dev_b));
    }
    template <unsigned Q, unsigned dof>
    void wallBoundary(double* dev_a, double* dev_a_post, double* dev_b = nullptr, double* dev_b_post = nullptr) {
        lattice<Q, dof>();
        check(clip_wall_boundary(m_s, dev_a, dev_a_post, dev_b, dev_b_post));
    }
    template <unsigned Q, unsigned dof>
    void slipWallBoundary(double* dev_a, double* dev_a_post, double* dev_b = nullptr, double* dev_b_post = nullptr) {
        lattice<Q, dof>();
        check(clip_slip_wall_boundary(m_s, dev_a, dev_a_post, dev_b, dev_b_post));
    }
    template <unsigned Q, unsigned dof>
    void freeConvectBoundary(double* dev_vel, double* dev_a, double* dev_a_prev, double* dev_b = nullptr,
                             double* dev_b_prev = nullptr) {
        lattice<Q, dof>();
        check(clip_free_convect_boundary(m_s, dev_vel, dev_a, dev_a_prev, dev_b, dev_b_prev));
    }
    template <unsigned Q, unsigned dof>
    void NeumannBoundary(double* dev_a, double* dev_b = nullptr) {
        lattice<Q, dof>();
        check(clip_neumann_boundary(m_s, dev_a, dev_b));
    }
    void velocityBoundary(double* dev_c, double* dev_f, double* dev_g) {
        check(clip_velocity_boundary(m_s, dev_c, dev_f, dev_g));
    }
    void mirrorBoundary(double* dev_a) { check(clip_mirror_boundary(m_s, dev_a)); }

    clip_solver* handle() const { return m_s; }

protected:
    template <unsigned Q, unsigned dof>
    void lattice() const {
        const bool ok = (m_idata->params.dim == 2) ? (Q == 9 && dof == 3) : (Q == 19 && dof == 5);
        if (!ok) throw std::runtime_error("clip: launcher template arguments do not match the lattice");
    }
    clip_solver* m_s = nullptr;
    const InputData* m_idata;
    const Domain* m_domain;
    const Boundary* m_boundary;
    const Geometry* m_geom;
};

class NSAllen : public Solver {
public:
    // The reference's constructor (NsAllen.cuh:50): the solver fills and reads
    // the DataArray's arrays at the points described there.
    NSAllen(const InputData& idata, const Domain& domain, DataArray& DA, const Boundary& boundary,
            const Geometry& geom, int precision = CLIP_PREC_F64)
        : Solver(idata, domain, boundary, geom, precision), m_DA(&DA) {
        DA.bind(this);
    }
    NSAllen(const InputData& idata, const Domain& domain, const Boundary& boundary, const Geometry& geom,
            int precision = CLIP_PREC_F64)
        : Solver(idata, domain, boundary, geom, precision) {}
    ~NSAllen() override {
        if (m_DA && m_DA->solver() == this) m_DA->bind(nullptr);
    }

    // NSAllen::initialCondition (NsAllen.cu:768-824): with a DataArray, its
    // host_c receives the initial phase field.
    void initialCondition() {
        m_pending_init = true;
        if (m_DA && m_DA->hostDA.host_c) check(clip_initial_phi(&m_idata->params, m_DA->hostDA.host_c));
    }
    // NSAllen::deviceInitializer (NsAllen.cu:897-908): the initial state from
    // the DataArray's dev_c (as updateDevice() left it), or from the initial
    // phase field without a DataArray.
    void deviceInitializer() {
        if (m_DA && m_DA->deviceDA.dev_c) {
            std::vector<double> phi(m_DA->domainSize());
            hip_check(hipMemcpy(phi.data(), m_DA->deviceDA.dev_c, phi.size() * sizeof(double), hipMemcpyDeviceToHost),
                      "NSAllen::deviceInitializer");
            check(clip_init_from_phi(m_s, phi.data()));
        } else {
            check(clip_init_fresh(m_s));
        }
        m_pending_init = false;
    }
    void solve(int n = 1) {
        m_resumed = false;
        check(clip_step(m_s, n));
    }
    // NSAllen::macroscopic (NsAllen.cu:866-877).  Right after
    // CheckPointer::load it is already done (the load leaves the handle as
    // load + macroscopic do, main.cu:67-71), so that call is skipped and a
    // loop written like main.cu resumes exactly.
    void macroscopic() {
        if (m_resumed) {
            m_resumed = false;
            return;
        }
        check(clip_macroscopic(m_s));
    }
    // Overwrites one of the device arrays (CLIP AoS layout, ghosts included):
    // CLIP_FIELD_PHI, U, P, F, G, F_POST or G_POST.
    void setField(int which, const std::vector<double>& values) {
        clip_info inf;
        check(clip_get_info(m_s, &inf));
        int nd = 1;
        if (which == CLIP_FIELD_U || which == CLIP_FIELD_DPHI || which == CLIP_FIELD_NORMAL) nd = inf.dim;
        if (which >= CLIP_FIELD_F) nd = inf.q;
        if (values.size() != inf.nodes * (size_t)nd) throw std::runtime_error("clip: setField size mismatch");
        m_resumed = false;
        check(clip_set_field(m_s, which, CLIP_LAYOUT_AOS, values.data()));
    }
    void markResumed() { m_resumed = true; }
    // Fields of the last macroscopic step, CLIP AoS layout, ghosts included.
    std::vector<double> field(int which) const {
        clip_info inf;
        check(clip_get_info(m_s, &inf));
        int nd = 1;
        if (which == CLIP_FIELD_U || which == CLIP_FIELD_DPHI || which == CLIP_FIELD_NORMAL) nd = inf.dim;
        if (which >= CLIP_FIELD_F) nd = inf.q;
        std::vector<double> out(inf.nodes * nd);
        check(clip_get_field(m_s, which, CLIP_LAYOUT_AOS, out.data()));
        return out;
    }
    DataArray* dataArray() const { return m_DA; }

private:
    DataArray* m_DA = nullptr;
    bool m_pending_init = false;
    bool m_resumed = false;
};

inline void DataArray::syncFromSolver() {
    if (!m_solver) throw std::runtime_error("clip: no solver is bound to this DataArray");
    if (!deviceDA.dev_c) throw std::runtime_error("clip: DataArray::createVectors was not called");
    const int Q = q(), D = m_params.dim;
    auto put = [&](int field, double* dst, int ndof, double* host_copy) {
        std::vector<double> v = m_solver->field(field);
        hip_check(hipMemcpy(dst, v.data(), m_nodes * ndof * sizeof(double), hipMemcpyHostToDevice),
                  "DataArray::syncFromSolver");
        if (host_copy) std::copy(v.begin(), v.end(), host_copy);
    };
    put(CLIP_FIELD_F, deviceDA.dev_f, Q, hostDA.host_f);
    put(CLIP_FIELD_G, deviceDA.dev_g, Q, hostDA.host_g);
    put(CLIP_FIELD_F_POST, deviceDA.dev_f_post, Q, hostDA.host_f_post);
    put(CLIP_FIELD_G_POST, deviceDA.dev_g_post, Q, hostDA.host_g_post);
    put(CLIP_FIELD_PHI, deviceDA.dev_c, 1, nullptr);
    put(CLIP_FIELD_RHO, deviceDA.dev_rho, 1, nullptr);
    put(CLIP_FIELD_P, deviceDA.dev_p, 1, nullptr);
    put(CLIP_FIELD_U, deviceDA.dev_vel, D, nullptr);
    put(CLIP_FIELD_MU, deviceDA.dev_mu, 1, nullptr);
    put(CLIP_FIELD_DPHI, deviceDA.dev_dc, D, nullptr);
    put(CLIP_FIELD_NORMAL, deviceDA.dev_normal, D, nullptr);
}

inline void DataArray::updateHost() {
    if (m_solver) syncFromSolver();
    auto get = [&](double* dst, const double* src, int ndof) {
        hip_check(hipMemcpy(dst, src, m_nodes * ndof * sizeof(double), hipMemcpyDeviceToHost),
                  "DataArray::updateHost");
    };
    get(hostDA.host_c, deviceDA.dev_c, 1);
    get(hostDA.host_p, deviceDA.dev_p, 1);
    get(hostDA.host_rho, deviceDA.dev_rho, 1);
    get(hostDA.host_vel, deviceDA.dev_vel, m_params.dim);
}

// The writers below take the solver directly, or the DataArray the solver
// was built on (the reference's signatures); the solver is looked up when
// they run.
inline NSAllen& solver_of(DataArray& DA) {
    if (!DA.solver()) throw std::runtime_error("clip: no solver is bound to this DataArray");
    return *DA.solver();
}

class Reporter {
public:
    Reporter(NSAllen& ns, const InputData& idata, const TimeInfo& ti)
        : m_ns(&ns), m_interval(idata.params.report_interval), m_ti(&ti) {}
    // Reporter(DataArray&, InputData&, Domain&, TimeInfo&) (Reporter.cuh:20).
    Reporter(DataArray& DA, const InputData& idata, const Domain&, const TimeInfo& ti)
        : m_da(&DA), m_interval(idata.params.report_interval), m_ti(&ti) {}
    // Reporter::print (Reporter.cu:124-144): Sum(c) every reportInterval and
    // a divergence error when phi is not finite.
    void print() {
        if (m_interval == 0 || m_ti->getCurrentStep() % m_interval != 0) return;
        double s = 0;
        int rc = clip_check_health(ns().handle(), &s);
        std::printf("Step        : %d\nSum(c)      : %.6g\n", m_ti->getCurrentStep(), s);
        if (rc == CLIP_ERR_DIVERGED) throw std::runtime_error("Solver has diverged.");
        check(rc);
    }

private:
    NSAllen& ns() const { return m_ns ? *m_ns : solver_of(*m_da); }
    NSAllen* m_ns = nullptr;
    DataArray* m_da = nullptr;
    unsigned m_interval;
    const TimeInfo* m_ti;
};

// VTSwriter (VTSwriter.cu:20-205): results/results_t<step>.vts every
// outputInterval steps; the fields leave the device only on those steps.
class VTSwriter {
public:
    VTSwriter(NSAllen& ns, const InputData& idata, const TimeInfo& ti, const std::string& folder = "results",
              const std::string& baseName = "results")
        : m_ns(&ns), m_interval(idata.params.output_interval), m_ti(&ti), m_folder(folder), m_base(baseName) {}
    // VTSwriter(DataArray&, InputData&, Domain&, TimeInfo&, folder, baseName) (VTSwriter.cuh:52).
    VTSwriter(DataArray& DA, const InputData& idata, const Domain&, const TimeInfo& ti,
              const std::string& folder = "results", const std::string& baseName = "results")
        : m_da(&DA), m_interval(idata.params.output_interval), m_ti(&ti), m_folder(folder), m_base(baseName) {}
    void writeToFile() {
        if (m_interval == 0 || m_ti->getCurrentStep() % m_interval != 0) return;
        std::filesystem::create_directories(m_folder);
        const std::string path = m_folder + "/" + m_base + "_t" + std::to_string(m_ti->getCurrentStep()) + ".vts";
        check(clip_write_vts(ns().handle(), path.c_str()));
    }

private:
    NSAllen& ns() const { return m_ns ? *m_ns : solver_of(*m_da); }
    NSAllen* m_ns = nullptr;
    DataArray* m_da = nullptr;
    unsigned m_interval;
    const TimeInfo* m_ti;
    std::string m_folder, m_base;
};

class CheckPointer {
public:
    CheckPointer(NSAllen& ns, const InputData& idata, TimeInfo& ti)
        : m_ns(&ns), m_copy(idata.params.checkpoint_copy), m_interval(idata.params.checkpoint_interval), m_ti(&ti) {}
    // CheckPointer(DataArray&, InputData&, Domain&, TimeInfo&, Boundary&) (CheckPointer.cuh:48).
    CheckPointer(DataArray& DA, const InputData& idata, const Domain&, TimeInfo& ti, const Boundary&)
        : m_da(&DA), m_copy(idata.params.checkpoint_copy), m_interval(idata.params.checkpoint_interval), m_ti(&ti) {}
    // CheckPointer::save (CheckPointer.cu:53-79): rotate, then write "Checkpoint".
    void save(const std::string& folder = "Checkpoint") {
        if (m_copy == 0 || m_interval == 0 || m_ti->getCurrentStep() % m_interval != 0) return;
        check(clip_checkpoint_rotate(folder.c_str(), m_copy));
        check(clip_checkpoint_save(ns().handle(), folder.c_str()));
    }
    // CheckPointer::load + NSAllen::macroscopic (CheckPointer.cu:84-102, main.cu:67-71).
    void load(const std::string& folder = "Checkpoint") {
        check(clip_checkpoint_load(ns().handle(), folder.c_str()));
        ns().markResumed();
        clip_info inf;
        check(clip_get_info(ns().handle(), &inf));
        m_ti->setCurrentStep((int)inf.step);
    }

private:
    NSAllen& ns() const { return m_ns ? *m_ns : solver_of(*m_da); }
    NSAllen* m_ns = nullptr;
    DataArray* m_da = nullptr;
    unsigned m_copy, m_interval;
    TimeInfo* m_ti;
};

}  // namespace clip
