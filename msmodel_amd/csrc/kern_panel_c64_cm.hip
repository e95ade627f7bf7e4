// kern_panel_c64_cm.hip -- the complex64, column-major instantiations of the panel kernel.
#define MSG_PANEL_IMPL
#include "kern_panel.h"

namespace msg {
template void launch_panel_rm<float, false>(cx<float>*, int, long long, int*, int*, int*, int, int, int, hipStream_t,
                                     int, int, int);
}  // namespace msg
