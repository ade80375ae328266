// Step-pass launchers for lines of 6, 12, 24, 48, 96 points (one of the per-length
// instantiation units, compiled in parallel).
#include "xpfc_lenops.hpp"

namespace xpfc {

XPFC_INSTANTIATE_LEN(6)
XPFC_INSTANTIATE_LEN(12)
XPFC_INSTANTIATE_LEN(24)
XPFC_INSTANTIATE_LEN(48)
XPFC_INSTANTIATE_LEN(96)

}  // namespace xpfc
