// kernels_mixed.hip -- mixed-precision interaction kernels (placeholder until
// the FP32 group-relative kernels land).
#include "internal.hpp"

namespace sphgpu {
static void not_ready() { throw Error(-6, "mixed-precision kernels are not built yet"); }
void launch_dens_mixed(const DensArgs &, hipStream_t) { not_ready(); }
void launch_drvt_mixed(const DrvtArgs &, hipStream_t) { not_ready(); }
void launch_hydro_mixed(const HydroArgs &, hipStream_t) { not_ready(); }
void launch_grav_mixed(const GravArgs &, hipStream_t) { not_ready(); }
}  // namespace sphgpu
