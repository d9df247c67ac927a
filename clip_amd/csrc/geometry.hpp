// geometry.hpp — signed distance functions, initial phase field and inlet
// profile (host side).
//
//   sdf            Geometry::sdf                  include/geometry/Geometry.cuh:102-161
//   initial phi    NSAllen::initialCondition      src/equation/NsAllen.cu:768-824
#pragma once

#include "clip_amd.h"

#include <cmath>

namespace clipamd {

// Signed distance to the first geometry object carrying `id` (positive
// outside); 1e10 when no object has that id.
inline double sdf(const clip_config& c, int id, double x, double y, double z) {
    for (int i = 0; i < c.num_geometries; ++i) {
        if (c.geom_id[i] != id) continue;
        const double* ce = c.geom_center[i];
        const double* le = c.geom_length[i];
        switch (c.geom_type[i]) {
            case CLIP_GEOM_CIRCLE: {
                const double dx = x - ce[0], dy = y - ce[1];
                return std::sqrt(dx * dx + dy * dy) - c.geom_radius[i];
            }
            case CLIP_GEOM_SPHERE: {
                const double dx = x - ce[0], dy = y - ce[1], dz = z - ce[2];
                return std::sqrt(dx * dx + dy * dy + dz * dz) - c.geom_radius[i];
            }
            case CLIP_GEOM_SQUARE:
            case CLIP_GEOM_CUBE: {
                const bool cube = c.geom_type[i] == CLIP_GEOM_CUBE;
                const double d[3] = {std::fabs(x - ce[0]) - 0.5 * le[0], std::fabs(y - ce[1]) - 0.5 * le[1],
                                     cube ? std::fabs(z - ce[2]) - 0.5 * le[2] : -1e300};
                double out2 = 0.0, inside = d[0];
                for (int a = 0; a < (cube ? 3 : 2); ++a) {
                    const double pos = d[a] > 0.0 ? d[a] : 0.0;
                    out2 += pos * pos;
                    if (d[a] > inside) inside = d[a];
                }
                return std::sqrt(out2) + (inside < 0.0 ? inside : 0.0);
            }
            case CLIP_GEOM_PERTURBATION: {
                const double kx = 2.0 * M_PI / le[0];
                const double amp = c.geom_amplitude[i] * le[0];
                const double shape = (c.dim == 2) ? std::cos(kx * x) : std::cos(kx * x) + std::cos(kx * z);
                return y - (amp * shape + ce[1]);
            }
            default:
                break;
        }
    }
    return 1e10;
}

// Phase field at lattice node (i, j, k) before the first step.
inline double initial_phi(const clip_config& c, int i, int j, int k) {
    const double z = (c.dim == 3) ? (double)k : 0.0;
    const double W = c.interface_width;
    switch (c.case_type) {
        case CLIP_CASE_BUBBLE:
        case CLIP_CASE_RTI:
            return 0.5 + 0.5 * std::tanh(2.0 * sdf(c, 0, i, j, z) / W);
        case CLIP_CASE_DROP:
        case CLIP_CASE_JET:
            return 0.5 - 0.5 * std::tanh(2.0 * sdf(c, 0, i, j, z) / W);
        default:
            return 0.0;
    }
}

// Inlet velocity of velocity side `side` at ghost node `pos`: the side's
// value, times the optional azimuthal modulation around geometry 0's centre.
inline void inlet_velocity(const clip_config& c, int side, const int pos[3], double U[3]) {
    for (int d = 0; d < 3; ++d) U[d] = c.bc_value[side][d];
    if (c.inlet_modes <= 0 || c.num_geometries <= 0) return;
    const int ax = side / 2;
    int t1 = (ax + 1) % 3, t2 = (ax + 2) % 3;
    if (c.dim == 2) {
        t1 = 1 - ax;
        t2 = -1;
    }
    const double d1 = pos[t1] - c.geom_center[0][t1];
    const double d2 = (t2 >= 0) ? pos[t2] - c.geom_center[0][t2] : 0.0;
    const double th = std::atan2(d2, d1);
    double mod = 1.0;
    for (int m = 0; m < c.inlet_modes; ++m) mod += c.inlet_mode_amp[m] * std::cos((m + 1) * th + c.inlet_mode_phase[m]);
    for (int d = 0; d < 3; ++d) U[d] *= mod;
}

}  // namespace clipamd
