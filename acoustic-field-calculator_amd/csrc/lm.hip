// Levenberg-Marquardt hologram solver (work in progress).
#include "afh_holo.hpp"
#include "afh_internal.hpp"

namespace afh {

bool lm_solve_single(const HoloProblem&, float*) {
    set_error("AFO_LM: not built yet");
    return false;
}

}  // namespace afh
