// Tile-pass kernel instances: five terms per pass, double, complex values.
#include "tbg_block_pass.h"

namespace tbg {

template void launch_pass_rt<double, 5, false>(int, const PassArgs&, const Blocking&,
                                                hipStream_t);
template void allow_lds_rt<double, 5, false>(int);

}  // namespace tbg
