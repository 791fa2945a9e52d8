// Error state and device queries of the pe_amd C ABI.
#include "pe_common.h"

#include <cstring>

namespace pe {

static thread_local std::string g_error;

void set_error(const std::string &message) { g_error = message; }
void clear_error() { g_error.clear(); }

int require_device(int *device) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
    (void)hipGetLastError();
    return fail(PE_ERR_NODEVICE, "no HIP device is visible");
  }
  int dev = 0;
  PE_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  PE_HIP(hipGetDeviceProperties(&prop, dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(PE_ERR_NODEVICE, std::string("device ") + prop.gcnArchName +
                                     " is not gfx950 (MI355X)");
  if (device) *device = dev;
  return PE_OK;
}

}  // namespace pe

extern "C" const char *pe_last_error(void) { return pe::g_error.c_str(); }

extern "C" int pe_abi_version(void) { return PE_AMD_ABI; }

extern "C" int pe_abi_struct_size(const char *name) {
  if (!name) return -1;
  const struct {
    const char *name;
    int size;
  } table[] = {
      {"pe_hji5d_config", (int)sizeof(pe_hji5d_config)},
      {"pe_hji5d_stats", (int)sizeof(pe_hji5d_stats)},
      {"pe_sl_car", (int)sizeof(pe_sl_car)},
      {"pe_sl_game", (int)sizeof(pe_sl_game)},
      {"pe_sl_stats", (int)sizeof(pe_sl_stats)},
      {"pe_mlp_config", (int)sizeof(pe_mlp_config)},
      {"pe_world8d_config", (int)sizeof(pe_world8d_config)},
      {"pe_world8d_stats", (int)sizeof(pe_world8d_stats)},
      {"pe_car_axes", (int)sizeof(pe_car_axes)},
  };
  for (const auto &t : table)
    if (std::strcmp(t.name, name) == 0) return t.size;
  return -1;
}

namespace pe {
__global__ __launch_bounds__(256) void copy_probe(const float4 *__restrict__ src,
                                                  float4 *__restrict__ dst,
                                                  long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = src[i];
}
}  // namespace pe

extern "C" int pe_copy_bandwidth(int64_t bytes, int32_t repeats, double *gbps) {
  pe::clear_error();
  if (bytes < 16 || bytes % 16 || repeats < 1 || !gbps)
    return pe::fail(PE_ERR_INVALID, "bytes must be a positive multiple of 16");
  int rc = pe::require_device(nullptr);
  if (rc) return rc;
  void *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto body = [&]() -> int {
    PE_HIP(hipMalloc(&a, bytes));
    PE_HIP(hipMalloc(&b, bytes));
    PE_HIP(hipMemset(a, 0, bytes));
    PE_HIP(hipEventCreate(&e0));
    PE_HIP(hipEventCreate(&e1));
    const long long n = bytes / 16;
    hipLaunchKernelGGL(pe::copy_probe, dim3(8192), dim3(256), 0, 0, (const float4 *)a,
                       (float4 *)b, n);
    PE_HIP(hipEventRecord(e0, 0));
    for (int r = 0; r < repeats; r++)
      hipLaunchKernelGGL(pe::copy_probe, dim3(8192), dim3(256), 0, 0, (const float4 *)a,
                         (float4 *)b, n);
    PE_HIP(hipEventRecord(e1, 0));
    PE_HIP(hipEventSynchronize(e1));
    float ms = 0;
    PE_HIP(hipEventElapsedTime(&ms, e0, e1));
    *gbps = 2.0 * bytes * repeats / (ms * 1e-3) / 1e9;
    return PE_OK;
  };
  rc = body();
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  return rc;
}

extern "C" int pe_device_info(char *name, int name_len, int32_t *compute_units,
                              int64_t *hbm_bytes) {
  int dev = 0;
  int rc = pe::require_device(&dev);
  if (rc) return rc;
  hipDeviceProp_t prop;
  PE_HIP(hipGetDeviceProperties(&prop, dev));
  if (name && name_len > 0) {
    std::snprintf(name, (size_t)name_len, "%s (%s)", prop.name,
                  prop.gcnArchName);
  }
  if (compute_units) *compute_units = prop.multiProcessorCount;
  if (hbm_bytes) *hbm_bytes = (int64_t)prop.totalGlobalMem;
  return PE_OK;
}
