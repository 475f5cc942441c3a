// common.cc — error plumbing and device metadata for the C-ABI.
#include "common.h"

#include <cstring>

namespace apgx {
namespace {
thread_local std::string g_last_error;
}  // namespace

int Fail(int code, const std::string& detail) {
  g_last_error = detail;
  return code;
}

void ClearError() { g_last_error.clear(); }

int HipFail(hipError_t error, const char* operation) {
  std::string detail = std::string(operation) + " failed: " + hipGetErrorString(error);
  if (error == hipErrorOutOfMemory || error == hipErrorMemoryAllocation) {
    return Fail(APGX_E_RESOURCE_EXHAUSTED, detail);
  }
  return Fail(APGX_E_BACKEND_FAILURE, detail);
}

}  // namespace apgx

extern "C" const char* apgx_last_error(void) { return apgx::g_last_error.c_str(); }

extern "C" uint32_t apgx_abi_version(void) { return (1u << 16) | 2u; }

extern "C" int apgx_device_synchronize(void) {
  APGX_HIP(hipDeviceSynchronize(), "hipDeviceSynchronize");
  apgx::ClearError();
  return APGX_OK;
}

// QueryMetadata (include/apgar/gpu/planar_router.h:550): name, architecture,
// runtime/driver versions and memory of the current device.
extern "C" int apgx_device_info_query(apgx_device_info* out) {
  if (out == nullptr) {
    return apgx::Fail(APGX_E_INTERNAL_INVARIANT, "apgx_device_info_query: null output");
  }
  std::memset(out, 0, sizeof(*out));
  int device = 0;
  APGX_HIP(hipGetDevice(&device), "hipGetDevice");
  hipDeviceProp_t prop;
  APGX_HIP(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
  std::strncpy(out->backend, "hip", sizeof(out->backend) - 1);
  std::strncpy(out->device_name, prop.name, sizeof(out->device_name) - 1);
  std::strncpy(out->gcn_arch_name, prop.gcnArchName, sizeof(out->gcn_arch_name) - 1);
  out->compute_units = static_cast<uint32_t>(prop.multiProcessorCount);
  int runtime = 0, driver = 0;
  APGX_HIP(hipRuntimeGetVersion(&runtime), "hipRuntimeGetVersion");
  APGX_HIP(hipDriverGetVersion(&driver), "hipDriverGetVersion");
  out->runtime_version = static_cast<uint32_t>(runtime);
  out->driver_version = static_cast<uint32_t>(driver);
  out->global_memory_bytes = static_cast<uint64_t>(prop.totalGlobalMem);
  apgx::ClearError();
  return APGX_OK;
}

// CudaCandidateDeviceClass (src/candidates/gpu_candidate_adapter.cc:27-35) for
// this backend: "hip-" + the device's gfx architecture name (the part before
// any ':' feature suffix), e.g. "hip-gfx950".  Incomplete metadata has no
// class, as there.  The marketing name is not required: some ROCm installs
// leave it empty, and the architecture is what identifies the class.
extern "C" int apgx_device_class(char* out, uint64_t capacity) {
  using namespace apgx;
  if (out == nullptr || capacity == 0) return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_device_class: null output");
  out[0] = '\0';
  apgx_device_info info;
  if (int s = apgx_device_info_query(&info)) return s;
  std::string arch(info.gcn_arch_name);
  arch = arch.substr(0, arch.find(':'));
  if (std::string(info.backend) != "hip" || arch.empty() || info.runtime_version == 0 || info.driver_version == 0 ||
      info.global_memory_bytes == 0)
    return Fail(APGX_E_UNSUPPORTED, "apgx_device_class: device metadata is incomplete");
  const std::string cls = "hip-" + arch;
  if (cls.size() + 1 > capacity) return Fail(APGX_E_RESOURCE_EXHAUSTED, "apgx_device_class: buffer too small");
  std::memcpy(out, cls.c_str(), cls.size() + 1);
  ClearError();
  return APGX_OK;
}
