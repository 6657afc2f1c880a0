// Temporary: families not yet implemented report hipErrorNotSupported.
#include "nnedi3.h"
#include "ravu.h"
namespace mpx {
int zoom_plan_create(mpx_ctx*, mpx_plan*, int, int, int, int) { return 0; }
void zoom_plan_destroy(mpx_ctx*, mpx_plan*) {}
hipError_t launch_ravu_3x(int, int, const Io&, const __half*, hipStream_t) { return hipErrorNotSupported; }
hipError_t launch_ravu(int, int, const Io&, const __half*, float*, int64_t, hipStream_t) { return hipErrorNotSupported; }
hipError_t launch_ravu_zoom(int, bool, int, const Io&, const __half*, const __half*, const ZoomPlan&, int, int,
                            hipStream_t) { return hipErrorNotSupported; }
int nnedi3_prepare(mpx_ctx*, int, int, int, NnediWeights*) { return 0; }
hipError_t launch_nnedi3(const NnediWeights&, const Io&, bool, bool, float*, int64_t, hipStream_t) {
  return hipErrorNotSupported;
}
}  // namespace mpx
