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
  };
  for (const auto &t : table)
    if (std::strcmp(t.name, name) == 0) return t.size;
  return -1;
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
