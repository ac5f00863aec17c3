// Process-wide settings of the ABI layer (set through ParODEExt.h).
#pragma once
#include <string>

namespace parode {

long& max_batch_steps();          // 0 = plan from free device memory
std::string& last_error();        // text of the last exception at the ABI (per thread)
int& level_kernel();              // bit 0: vector FP32 levels, bit 1: vector FP64 GEMM, bit 2: no fused top,
                                  // bit 3: no ABM chunk recurrences

}  // namespace parode
