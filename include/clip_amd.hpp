// clip_amd.hpp — C++ host surface over the C-ABI, with the reference's class
// names so that code written against CLIP's host classes maps one to one:
//
//   clip::InputData     include/input/InputData.cuh:42-151   (config.txt -> SimParams)
//   clip::Domain        include/domain/Domain.cuh:30-112      (extents, ghost ring)
//   clip::Boundary      include/boundary/Boundary.cuh:48-190  (BC map and flags)
//   clip::Geometry      include/geometry/Geometry.cuh         (geometry pool)
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

class NSAllen {
public:
    NSAllen(const InputData& idata, const Domain&, const Boundary&, const Geometry&, int precision = CLIP_PREC_F64) {
        check(clip_create(&idata.params, precision, &m_s));
    }
    ~NSAllen() { clip_destroy(m_s); }
    NSAllen(const NSAllen&) = delete;
    NSAllen& operator=(const NSAllen&) = delete;

    // initialCondition + updateDevice + deviceInitializer (NsAllen.cu:768-908)
    void initialCondition() { m_pending_init = true; }
    void deviceInitializer() {
        if (m_pending_init) check(clip_init_fresh(m_s));
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
    clip_solver* handle() const { return m_s; }

private:
    clip_solver* m_s = nullptr;
    bool m_pending_init = false;
    bool m_resumed = false;
};

class Reporter {
public:
    Reporter(NSAllen& ns, const InputData& idata, const TimeInfo& ti)
        : m_ns(&ns), m_interval(idata.params.report_interval), m_ti(&ti) {}
    // Reporter::print (Reporter.cu:124-144): Sum(c) every reportInterval and
    // a divergence error when phi is not finite.
    void print() {
        if (m_interval == 0 || m_ti->getCurrentStep() % m_interval != 0) return;
        double s = 0;
        int rc = clip_check_health(m_ns->handle(), &s);
        std::printf("Step        : %d\nSum(c)      : %.6g\n", m_ti->getCurrentStep(), s);
        if (rc == CLIP_ERR_DIVERGED) throw std::runtime_error("Solver has diverged.");
        check(rc);
    }

private:
    NSAllen* m_ns;
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
    void writeToFile() {
        if (m_interval == 0 || m_ti->getCurrentStep() % m_interval != 0) return;
        std::filesystem::create_directories(m_folder);
        const std::string path = m_folder + "/" + m_base + "_t" + std::to_string(m_ti->getCurrentStep()) + ".vts";
        check(clip_write_vts(m_ns->handle(), path.c_str()));
    }

private:
    NSAllen* m_ns;
    unsigned m_interval;
    const TimeInfo* m_ti;
    std::string m_folder, m_base;
};

class CheckPointer {
public:
    CheckPointer(NSAllen& ns, const InputData& idata, TimeInfo& ti)
        : m_ns(&ns), m_copy(idata.params.checkpoint_copy), m_interval(idata.params.checkpoint_interval), m_ti(&ti) {}
    // CheckPointer::save (CheckPointer.cu:53-79): rotate, then write "Checkpoint".
    void save(const std::string& folder = "Checkpoint") {
        if (m_copy == 0 || m_interval == 0 || m_ti->getCurrentStep() % m_interval != 0) return;
        check(clip_checkpoint_rotate(folder.c_str(), m_copy));
        check(clip_checkpoint_save(m_ns->handle(), folder.c_str()));
    }
    // CheckPointer::load + NSAllen::macroscopic (CheckPointer.cu:84-102, main.cu:67-71).
    void load(const std::string& folder = "Checkpoint") {
        check(clip_checkpoint_load(m_ns->handle(), folder.c_str()));
        m_ns->markResumed();
        clip_info inf;
        check(clip_get_info(m_ns->handle(), &inf));
        m_ti->setCurrentStep((int)inf.step);
    }

private:
    NSAllen* m_ns;
    unsigned m_copy, m_interval;
    TimeInfo* m_ti;
};

}  // namespace clip
