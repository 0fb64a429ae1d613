// Library lifecycle: ABI version, thread-local error messages, device query.
#include "fmc_common.h"

#include <string.h>

namespace fmc {

static thread_local char g_last_error[512] = {0};

void set_error(int code, const char* fmt, ...) {
  int n = snprintf(g_last_error, sizeof(g_last_error), "[fmc status %d] ", code);
  if (n < 0 || n >= (int)sizeof(g_last_error)) n = 0;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error + n, sizeof(g_last_error) - n, fmt, ap);
  va_end(ap);
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(FMC_ERR_LAUNCH, "%s: %s", what, hipGetErrorString(e));
    return FMC_ERR_LAUNCH;
  }
  return FMC_OK;
}

}  // namespace fmc

extern "C" int fmc_abi_version(void) { return FMC_ABI_VERSION; }

extern "C" const char* fmc_last_error(void) { return fmc::g_last_error; }

extern "C" int fmc_device_query(int device, int* n_cu, int* lds_bytes_per_block, int* warp_size, int* arch_major,
                                int* arch_minor) {
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) {
    fmc::set_error(FMC_ERR_HIP, "hipGetDeviceProperties(%d): %s", device, hipGetErrorString(e));
    return FMC_ERR_HIP;
  }
  if (n_cu) *n_cu = prop.multiProcessorCount;
  if (lds_bytes_per_block) *lds_bytes_per_block = (int)prop.sharedMemPerBlock;
  if (warp_size) *warp_size = prop.warpSize;
  if (arch_major) *arch_major = prop.major;
  if (arch_minor) *arch_minor = prop.minor;
  return FMC_OK;
}
