// Step-pass launchers for lines of 128, 256 points (one of the per-length
// instantiation units, compiled in parallel).
#include "xpfc_lenops.hpp"

namespace xpfc {

XPFC_INSTANTIATE_LEN(128)
XPFC_INSTANTIATE_LEN(256)

}  // namespace xpfc
