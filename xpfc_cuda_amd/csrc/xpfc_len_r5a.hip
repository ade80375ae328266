// Step-pass launchers for lines of 10, 20, 40, 80 points (one of the per-length
// instantiation units, compiled in parallel).
#include "xpfc_lenops.hpp"

namespace xpfc {

XPFC_INSTANTIATE_LEN(10)
XPFC_INSTANTIATE_LEN(20)
XPFC_INSTANTIATE_LEN(40)
XPFC_INSTANTIATE_LEN(80)

}  // namespace xpfc
