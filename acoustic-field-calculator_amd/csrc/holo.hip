// Hologram solvers on the GPU (work in progress).
#include "afh_holo.hpp"
#include "afh_internal.hpp"

namespace afh {

bool holo_solve_single(const HoloProblem&, int, float*, float*) {
    set_error("hologram solvers: not built yet");
    return false;
}

}  // namespace afh
