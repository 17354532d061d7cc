// Status strings, ABI version and device information.
#include "adt_common.hpp"

#include <cstring>

namespace adt {

static thread_local char g_error[512] = "";

int set_error(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
    return code;
}

}  // namespace adt

extern "C" const char* adt_last_error(void) { return adt::g_error; }

extern "C" int adt_abi_version(void) { return 7; }

extern "C" int adt_device_info(int device, int* h_cus, char* h_name, int name_len) {
    hipDeviceProp_t prop;
    ADT_HIP_CHECK(hipGetDeviceProperties(&prop, device));
    if (h_cus) *h_cus = prop.multiProcessorCount;
    if (h_name && name_len > 0) {
        std::strncpy(h_name, prop.gcnArchName, static_cast<size_t>(name_len - 1));
        h_name[name_len - 1] = 0;
    }
    return ADT_OK;
}
