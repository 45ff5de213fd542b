// Tile-pass kernel instances: two terms per pass (the pass that ends a step).
#include "tbg_block_pass.h"

namespace tbg {

template void launch_pass_rt<float, 2, true>(int, const PassArgs&, const Blocking&,
                                                hipStream_t);
template void allow_lds_rt<float, 2, true>(int);
template void launch_pass_rt<float, 2, false>(int, const PassArgs&, const Blocking&,
                                                hipStream_t);
template void allow_lds_rt<float, 2, false>(int);
template void launch_pass_rt<double, 2, true>(int, const PassArgs&, const Blocking&,
                                                hipStream_t);
template void allow_lds_rt<double, 2, true>(int);
template void launch_pass_rt<double, 2, false>(int, const PassArgs&, const Blocking&,
                                                hipStream_t);
template void allow_lds_rt<double, 2, false>(int);

}  // namespace tbg
