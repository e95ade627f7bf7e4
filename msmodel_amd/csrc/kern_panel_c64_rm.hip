// kern_panel_c64_rm.hip -- the complex64, row-major instantiations of the panel kernel.
#define MSG_PANEL_IMPL
#include "kern_panel.h"

namespace msg {
template void launch_panel_rm<float, true>(cx<float>*, int, long long, int*, int*, int*, int, int, int, hipStream_t,
                                     int, int, int);
}  // namespace msg
