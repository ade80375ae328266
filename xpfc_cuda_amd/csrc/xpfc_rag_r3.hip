// Ragged (several-grid) step passes over lines of the 3*2^k-point family.
#include "xpfc_ragged.hpp"

namespace xpfc {
XPFC_INSTANTIATE_RAG(3)
}  // namespace xpfc
