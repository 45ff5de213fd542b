// Tile-pass kernel instances: four terms per pass, float, real values.
#include "tbg_block_pass.h"

namespace tbg {

template void launch_pass_rt<float, 4, true>(int, const PassArgs&, const Blocking&,
                                                hipStream_t);
template void allow_lds_rt<float, 4, true>(int);

}  // namespace tbg
