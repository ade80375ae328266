// Ragged (several-grid) step passes over lines of the 2^k-point family.
#include "xpfc_ragged.hpp"

namespace xpfc {
XPFC_INSTANTIATE_RAG(1)
}  // namespace xpfc
