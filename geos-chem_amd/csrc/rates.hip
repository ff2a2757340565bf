// Update_RCONST on the GPU: placeholder until the generated rate program
// lands; fails loudly rather than producing rates.
#include <hip/hip_runtime.h>

#include "gckpp_internal.h"

namespace gckpp {
bool rates_available(int) { return false; }
void launch_rates_mech(int, const RatesArgs &, hipStream_t) {}
}  // namespace gckpp
