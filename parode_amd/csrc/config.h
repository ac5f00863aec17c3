// Process-wide settings of the ABI layer (set through ParODEExt.h).
#pragma once
#include <string>

namespace parode {

long& max_batch_steps();          // 0 = plan from free device memory
std::string& last_error();        // text of the last exception at the ABI
int& level_kernel();              // FP32 levels: 0 = auto (MFMA when the size allows), 1 = SIMT

}  // namespace parode
