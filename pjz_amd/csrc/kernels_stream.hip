// Streaming stepper: placeholder until the fused x-sweep kernel lands.
#include "fdtd_common.h"

namespace pjz {

hipError_t run_stream(const Geom&, const pjz_fdtd_problem&, const float*, const Src&,
                      const OutWin&, char*, const Layout&, int, hipStream_t) {
  return hipErrorNotSupported;
}

}  // namespace pjz
