#include "ufunc.h"

#include <hip/hiprtc.h>

#include <algorithm>

#include "hip_util.h"

namespace parode {

namespace {
// The user source is compiled as host+device code so that the OpenCL-C style
// function (no __device__ qualifier) is callable from the sampler.
const char* kPrelude =
    "typedef float fType;\n"
    "#pragma clang force_cuda_host_device begin\n";
const char* kSampler =
    "\n#pragma clang force_cuda_host_device end\n"
    "extern \"C\" __global__ void parode_u_sample(float* __restrict__ u, float t0, float dt,\n"
    "                                             long nSamples, int nInputs)\n"
    "{\n"
    "    long total = nSamples * nInputs;\n"
    "    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;\n"
    "         idx += (long)gridDim.x * blockDim.x) {\n"
    "        long s = idx / nInputs;\n"
    "        int j = (int)(idx % nInputs);\n"
    "        float t = __builtin_fmaf(dt, (float)s, t0);\n"
    "        u[idx] = (float)UFunc(t, j);\n"
    "    }\n"
    "}\n";
}  // namespace

bool UKernel::compile(const std::string& src, const std::string& arch, std::vector<char>& code,
                      std::string& log)
{
    std::string full = std::string(kPrelude) + src + kSampler;
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, full.c_str(), "parode_ufunc.hip", 0, nullptr, nullptr) !=
        HIPRTC_SUCCESS) {
        log = "hiprtcCreateProgram failed";
        return false;
    }
    std::string archOpt = "--offload-arch=" + arch;
    const char* opts[] = {archOpt.c_str(), "-O3"};
    hiprtcResult r = hiprtcCompileProgram(prog, 2, opts);
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    log.assign(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &log[0]);
    bool ok = r == HIPRTC_SUCCESS;
    if (ok) {
        size_t cs = 0;
        hiprtcGetCodeSize(prog, &cs);
        code.resize(cs);
        hiprtcGetCode(prog, code.data());
        ok = cs > 0;
    }
    hiprtcDestroyProgram(&prog);
    return ok;
}

UKernel::UKernel(const std::vector<char>& code)
{
    PARODE_HIP(hipModuleLoadData(&module_, code.data()));
    hipError_t e = hipModuleGetFunction(&fn_, module_, "parode_u_sample");
    if (e != hipSuccess) {
        (void)hipModuleUnload(module_);
        module_ = nullptr;
        throw HipError(e, "hipModuleGetFunction(parode_u_sample)");
    }
}

UKernel::~UKernel()
{
    if (module_) (void)hipModuleUnload(module_);
}

void UKernel::sample(hipStream_t s, float* dU, float t0, float dt, long nSamples, int m) const
{
    long total = nSamples * m;
    if (total <= 0) return;
    unsigned blocks = (unsigned)std::min<long>((total + 255) / 256, 65536);
    float* u = dU;
    long n = nSamples;
    int mm = m;
    void* params[] = {&u, &t0, &dt, &n, &mm};
    PARODE_HIP(hipModuleLaunchKernel(fn_, blocks, 1, 1, 256, 1, 1, 0, s, params, nullptr));
}

}  // namespace parode
