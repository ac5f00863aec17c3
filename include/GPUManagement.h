/* HIP device layer: the replacement of the reference's OpenCL runtime layer
 * (Include/GPUManagement.h, Sources/GPUManagement.cpp): device enumeration,
 * names, one stream per selected device, device buffers, the GPU timer. */
#ifndef PARODE_GPUMANAGEMENT_H_
#define PARODE_GPUMANAGEMENT_H_

#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>
#include <vector>

/* A device allocation, owning or a non-owning view (the role of
 * OpenCLMemBuffer, GPUManagement.h:130-160 of the reference). */
class DeviceBuffer {
public:
	DeviceBuffer() {}
	DeviceBuffer(std::size_t bytes, int device);                 /* owning */
	DeviceBuffer(void* view, std::size_t bytes, int device)      /* view */
		: _p(view), _bytes(bytes), _device(device), _owner(false) {}
	~DeviceBuffer();
	DeviceBuffer(const DeviceBuffer&) = delete;
	DeviceBuffer& operator=(const DeviceBuffer&) = delete;
	void* Get() const { return _p; }
	std::size_t Size() const { return _bytes; }
	int Device() const { return _device; }
	/* blocking copies on the given stream (offsets in bytes) */
	void MemWrite(const void* host, hipStream_t s, std::size_t offset = 0, std::size_t bytes = 0);
	void MemRead(void* host, hipStream_t s, std::size_t offset = 0, std::size_t bytes = 0) const;

private:
	void* _p = nullptr;
	std::size_t _bytes = 0;
	int _device = 0;
	bool _owner = false;
};

class GPUManagement {
public:
	GPUManagement();
	~GPUManagement();
	/* every visible device id (GPUManagement.cpp:64-105) */
	void GetGPUDevices(std::vector<unsigned int>& ids);
	/* initialise all devices / a subset; false when there is no GPU */
	bool Initialize(int nQueues = 1);
	bool Initialize(const std::vector<unsigned int>& selected, int nQueues = 1);
	int GetNumberOfDevices() const { return (int)_devices.size(); }
	std::string GetDeviceName(unsigned int id);
	int DeviceId(int index) const { return _devices[index]; }
	hipStream_t Stream(int index) const { return _streams[index]; }
	std::string Arch(int index) const;
	std::size_t GlobalMemorySize(int index) const;
	std::size_t FreeMemory(int index) const;
	int ComputeUnits(int index) const;
	void SetCurrent(int index) const;
	/* GPU timer on device 0's stream, milliseconds (GPUManagement.cpp:216-268) */
	void StartTimer();
	float StopTimer();

private:
	void Release();
	std::vector<int> _devices;
	std::vector<hipStream_t> _streams;
	std::vector<hipDeviceProp_t> _props;
	hipEvent_t _t0 = nullptr, _t1 = nullptr;
	bool _timerStarted = false;
};

#endif /* PARODE_GPUMANAGEMENT_H_ */
