// inst_f32.hip — FP32 instantiations.
#include "solver_impl.hpp"

namespace clipamd {

std::unique_ptr<Impl> make_impl_f32(const clip_config& c) {
    if (c.dim == 2) return std::make_unique<SolverImpl<D2Q9, float, float, StorePlain<float, float>>>(c);
    return std::make_unique<SolverImpl<D3Q19, float, float, StorePlain<float, float>>>(c);
}

}  // namespace clipamd
