// Tile-pass kernel instances: five terms per pass, float, real values.
#include "tbg_block_pass.h"

namespace tbg {

template void launch_pass_rt<float, 5, true>(int, const PassArgs&, const Blocking&,
                                                hipStream_t);
template void allow_lds_rt<float, 5, true>(int);

}  // namespace tbg
