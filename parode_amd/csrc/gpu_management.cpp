// HIP device layer (replaces the reference's OpenCL runtime layer,
// Sources/GPUManagement.cpp).
#include <GPUManagement.h>

#include <cstring>

#include "hip_util.h"

DeviceBuffer::DeviceBuffer(std::size_t bytes, int device) : _bytes(bytes), _device(device), _owner(true)
{
    if (bytes == 0) bytes = 1;   // the reference also never allocates 0 bytes (:81, :96)
    PARODE_HIP(hipSetDevice(device));
    PARODE_HIP(hipMalloc(&_p, bytes));
}

DeviceBuffer::~DeviceBuffer()
{
    if (_owner && _p) (void)hipFree(_p);
}

void DeviceBuffer::MemWrite(const void* host, hipStream_t s, std::size_t offset, std::size_t bytes)
{
    if (bytes == 0) bytes = _bytes - offset;
    if (offset + bytes > _bytes) throw std::out_of_range("DeviceBuffer::MemWrite");
    if (bytes == 0) return;
    PARODE_HIP(hipMemcpyAsync((char*)_p + offset, host, bytes, hipMemcpyHostToDevice, s));
    PARODE_HIP(hipStreamSynchronize(s));
}

void DeviceBuffer::MemRead(void* host, hipStream_t s, std::size_t offset, std::size_t bytes) const
{
    if (bytes == 0) bytes = _bytes - offset;
    if (offset + bytes > _bytes) throw std::out_of_range("DeviceBuffer::MemRead");
    if (bytes == 0) return;
    PARODE_HIP(hipMemcpyAsync(host, (const char*)_p + offset, bytes, hipMemcpyDeviceToHost, s));
    PARODE_HIP(hipStreamSynchronize(s));
}

GPUManagement::GPUManagement() {}

GPUManagement::~GPUManagement() { Release(); }

void GPUManagement::Release()
{
    for (size_t i = 0; i < _streams.size(); ++i) {
        (void)hipSetDevice(_devices[i]);
        (void)hipStreamSynchronize(_streams[i]);
        (void)hipStreamDestroy(_streams[i]);
    }
    if (_t0) (void)hipEventDestroy(_t0);
    if (_t1) (void)hipEventDestroy(_t1);
    _t0 = _t1 = nullptr;
    _streams.clear();
    _devices.clear();
    _props.clear();
}

void GPUManagement::GetGPUDevices(std::vector<unsigned int>& ids)
{
    ids.clear();
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) return;
    for (int i = 0; i < count; ++i) ids.push_back((unsigned)i);
}

bool GPUManagement::Initialize(int nQueues)
{
    std::vector<unsigned int> ids;
    GetGPUDevices(ids);
    return Initialize(ids, nQueues);
}

bool GPUManagement::Initialize(const std::vector<unsigned int>& selected, int)
{
    Release();
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return false;
    for (unsigned id : selected) {
        if ((int)id >= count) throw std::invalid_argument("invalid device id");
        hipDeviceProp_t prop;
        PARODE_HIP(hipGetDeviceProperties(&prop, (int)id));
        PARODE_HIP(hipSetDevice((int)id));
        hipStream_t s;
        PARODE_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        _devices.push_back((int)id);
        _streams.push_back(s);
        _props.push_back(prop);
    }
    if (_devices.empty()) return false;
    PARODE_HIP(hipSetDevice(_devices[0]));
    PARODE_HIP(hipEventCreate(&_t0));
    PARODE_HIP(hipEventCreate(&_t1));
    return true;
}

std::string GPUManagement::GetDeviceName(unsigned int id)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || (int)id >= count) return std::string();
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, (int)id) != hipSuccess) return std::string();
    return std::string(prop.name);
}

std::string GPUManagement::Arch(int index) const
{
    std::string a(_props[index].gcnArchName);
    return a.substr(0, a.find(':'));
}

std::size_t GPUManagement::GlobalMemorySize(int index) const { return _props[index].totalGlobalMem; }

std::size_t GPUManagement::FreeMemory(int index) const
{
    SetCurrent(index);
    size_t fr = 0, tot = 0;
    PARODE_HIP(hipMemGetInfo(&fr, &tot));
    return fr;
}

int GPUManagement::ComputeUnits(int index) const { return _props[index].multiProcessorCount; }

void GPUManagement::SetCurrent(int index) const { PARODE_HIP(hipSetDevice(_devices[index])); }

void GPUManagement::StartTimer()
{
    SetCurrent(0);
    PARODE_HIP(hipEventRecord(_t0, _streams[0]));
    _timerStarted = true;
}

float GPUManagement::StopTimer()
{
    SetCurrent(0);
    if (!_timerStarted) throw std::logic_error("StopTimer without StartTimer");
    PARODE_HIP(hipEventRecord(_t1, _streams[0]));
    PARODE_HIP(hipEventSynchronize(_t1));
    float ms = 0.0f;
    PARODE_HIP(hipEventElapsedTime(&ms, _t0, _t1));
    return ms;
}
