// Lyndon-grammar BWT (Lyndon_Grammar_BWT.cu).  Placeholder until the GPU pipeline lands.
#include "prims.hpp"

namespace mibwt {

int64_t lyndon_bwt_run(Storage& S, const uint8_t* src, bool src_host, uint8_t* dO, int64_t n)
{
    (void)S; (void)src; (void)src_host; (void)dO; (void)n;
    throw Error(LIBCUBWT_GPU_NOT_SUPPORTED, "Lyndon-grammar BWT is not built yet");
}

}  // namespace mibwt
