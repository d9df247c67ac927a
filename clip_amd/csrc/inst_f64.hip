// inst_f64.hip — FP64 instantiations (drop-in precision).  Compiled with
// -ffp-contract=off: every FP64 operation is rounded on its own, in the
// expression order the oracle uses, so the two agree bit for bit.
#include "solver_impl.hpp"

namespace clipamd {

std::unique_ptr<Impl> make_impl_f64(const clip_config& c) {
    if (c.dim == 2) return std::make_unique<SolverImpl<D2Q9, double, double, StorePlain<double, double>>>(c);
    return std::make_unique<SolverImpl<D3Q19, double, double, StorePlain<double, double>>>(c);
}

}  // namespace clipamd
