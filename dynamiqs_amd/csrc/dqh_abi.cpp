// Host-side pieces of the dqh C-ABI that carry no kernels.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "dqh.h"
#include "dqh_common.h"

namespace dqh {
static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }
int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}
}  // namespace dqh

extern "C" const char *dqh_last_error(void) { return dqh::g_last_error.c_str(); }

extern "C" int dqh_device_info(char *name, int64_t len, int64_t *cus) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return dqh::fail(DQH_HIP_ERROR, hipGetErrorString(e));
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, dev);
  if (e != hipSuccess) return dqh::fail(DQH_HIP_ERROR, hipGetErrorString(e));
  if (name && len > 0) {
    std::strncpy(name, prop.gcnArchName, (size_t)len - 1);
    name[len - 1] = 0;
  }
  if (cus) *cus = prop.multiProcessorCount;
  return DQH_OK;
}

extern "C" int dqh_host_device_pointer(void *host, void **device) {
  if (!host || !device) return dqh::fail(DQH_BAD_ARG, "dqh_host_device_pointer: null argument");
  void *d = nullptr;
  const hipError_t e = hipHostGetDevicePointer(&d, host, 0);
  if (e != hipSuccess || !d)
    return dqh::fail(DQH_HIP_ERROR, std::string("dqh_host_device_pointer: ") + hipGetErrorString(e));
  *device = d;
  return DQH_OK;
}

extern "C" int dqh_host_alloc(int64_t bytes, int64_t coherent, void **host) {
  if (!host || bytes <= 0) return dqh::fail(DQH_BAD_ARG, "dqh_host_alloc: bad arguments");
  void *h = nullptr;
  const unsigned flags = hipHostMallocMapped | (coherent ? hipHostMallocCoherent : hipHostMallocNonCoherent);
  const hipError_t e = hipHostMalloc(&h, (size_t)bytes, flags);
  if (e != hipSuccess || !h) return dqh::fail(DQH_HIP_ERROR, std::string("dqh_host_alloc: ") + hipGetErrorString(e));
  *host = h;
  return DQH_OK;
}

extern "C" int dqh_host_free(void *host) {
  if (!host) return DQH_OK;
  const hipError_t e = hipHostFree(host);
  return e == hipSuccess ? DQH_OK : dqh::fail(DQH_HIP_ERROR, std::string("dqh_host_free: ") + hipGetErrorString(e));
}
