// Step-pass launchers for lines of 160, 320, 640, 1280 points (one of the per-length
// instantiation units, compiled in parallel).
#include "xpfc_lenops.hpp"

namespace xpfc {

XPFC_INSTANTIATE_LEN(160)
XPFC_INSTANTIATE_LEN(320)
XPFC_INSTANTIATE_LEN(640)
XPFC_INSTANTIATE_LEN(1280)

}  // namespace xpfc
