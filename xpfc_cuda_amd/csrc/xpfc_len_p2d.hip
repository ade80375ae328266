// Step-pass launchers for lines of 2048 points (one of the per-length
// instantiation units, compiled in parallel).
#include "xpfc_lenops.hpp"

namespace xpfc {

XPFC_INSTANTIATE_LEN(2048)

}  // namespace xpfc
