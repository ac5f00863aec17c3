// Input functions: `fType UFunc(fType t, int nInputIndex)` source text, compiled
// at registration with hiprtc into a sampler kernel (the reference builds an
// OpenCL program from "#include <FTypeDef.h>" + UFunc + Kernels/u.cl,
// Sources/ParODE.cpp:410-466).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace parode {

class UKernel {
public:
    // Compiles src for a gfx target; returns true on success; log receives the
    // compiler output.  Needs no GPU.
    static bool compile(const std::string& src, const std::string& arch, std::vector<char>& code,
                        std::string& log);
    // Loads a compiled sampler on the current device.
    explicit UKernel(const std::vector<char>& code);
    ~UKernel();
    UKernel(const UKernel&) = delete;
    UKernel& operator=(const UKernel&) = delete;

    // u[s*m + j] = UFunc(t_s, j) with the float time grid of Kernels/u.cl:8,
    // t_s = t0 + dt * s (multiply-add contracted), s < nSamples.
    void sample(hipStream_t s, float* dU, float t0, float dt, long nSamples, int m) const;

private:
    hipModule_t module_ = nullptr;
    hipFunction_t fn_ = nullptr;
};

}  // namespace parode
