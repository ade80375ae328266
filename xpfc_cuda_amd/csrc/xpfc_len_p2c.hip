// Step-pass launchers for lines of 512, 1024 points (one of the per-length
// instantiation units, compiled in parallel).
#include "xpfc_lenops.hpp"

namespace xpfc {

XPFC_INSTANTIATE_LEN(512)
XPFC_INSTANTIATE_LEN(1024)

}  // namespace xpfc
