// kern_panel_c128_cm.hip -- the complex128, column-major instantiations of the panel kernel.
#define MSG_PANEL_IMPL
#include "kern_panel.h"

namespace msg {
template void launch_panel_rm<double, false>(cx<double>*, int, long long, int*, int*, int*, int, int, int, hipStream_t,
                                     int, int, int);
}  // namespace msg
