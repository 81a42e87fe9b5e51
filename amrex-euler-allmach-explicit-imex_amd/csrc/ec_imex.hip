// ec_imex.hip -- IMEX T1/S2 BDLTV20 source-map Picard step (device).
#include "ec_context.hpp"

namespace ec {
void imex_release(ec_ctx*) {}
}  // namespace ec
