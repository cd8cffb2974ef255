// Fused-step instantiations for one lattice (kept in its own translation unit so the
// lattices compile in parallel).
#include "step.hpp"

namespace xlbh {
int launch_step_d2q9(const StepCall& c) { return launch_step_lat<XLBH_D2Q9>(c); }
}  // namespace xlbh
