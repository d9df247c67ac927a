// inst_mixed.hip — mixed-precision instantiations: FP32 arithmetic on
// populations stored in 16 bits (FP16 or BF16) as scaled deviations from the
// lattice-weight equilibrium at rest; phi, u, p and the outflow planes in FP32,
// pressure summed and conservation sums reduced in FP64.
#include "solver_impl.hpp"

namespace clipamd {

std::unique_ptr<Impl> make_impl_mixed(const clip_config& c, bool bf16) {
    if (c.dim == 2) {
        if (bf16) return std::make_unique<SolverImpl<D2Q9, __hip_bfloat16, float, StoreMixed<D2Q9, __hip_bfloat16>>>(c);
        return std::make_unique<SolverImpl<D2Q9, __half, float, StoreMixed<D2Q9, __half>>>(c);
    }
    if (bf16) return std::make_unique<SolverImpl<D3Q19, __hip_bfloat16, float, StoreMixed<D3Q19, __hip_bfloat16>>>(c);
    return std::make_unique<SolverImpl<D3Q19, __half, float, StoreMixed<D3Q19, __half>>>(c);
}

}  // namespace clipamd
