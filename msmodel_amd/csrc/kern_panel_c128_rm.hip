// kern_panel_c128_rm.hip -- the complex128, row-major instantiations of the panel kernel.
#define MSG_PANEL_IMPL
#include "kern_panel.h"

namespace msg {
template void launch_panel_rm<double, true>(cx<double>*, int, long long, int*, int*, int*, int, int, int, hipStream_t,
                                     int, int, int);
}  // namespace msg
