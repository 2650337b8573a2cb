// (W,P) selection for libcubwt_bwt.  Fixed W=5,P=16 (the reference default,
// Benchmark_Skript_Linux.sh:70-71) until a trained selector is installed.
#include <cstdint>

namespace mibwt {

void selector_choose(const uint8_t* T, int64_t n, int* W, int* P)
{
    (void)T;
    (void)n;
    *W = 5;
    *P = 16;
}

}  // namespace mibwt
