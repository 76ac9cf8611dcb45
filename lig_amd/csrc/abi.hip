// Status names and version of the C-ABI.
#include "lig_device.h"

extern "C" const char *lig_strerror(int status) {
    switch (status) {
        case LIG_OK: return "ok";
        case LIG_ERR_INVALID_ARGUMENT: return "invalid argument";
        case LIG_ERR_WORKSPACE: return "workspace missing or too small";
        case LIG_ERR_LAUNCH: return "kernel launch failed";
        case LIG_ERR_UNSUPPORTED: return "unsupported configuration";
        case LIG_ERR_TOO_LARGE: return "input exceeds the entry point's fixed scratch size";
        default: return "unknown status";
    }
}

extern "C" int lig_abi_version(void) { return 100; }

namespace lig {
namespace {
constexpr int kMaxCheckUnits = 32;
struct CheckUnit {
    const char *file;
    CheckReader read;
};
CheckUnit g_check_units[kMaxCheckUnits];  // filled by static initialisers
int g_check_count = 0;
}  // namespace

int register_check_reader(const char *file, CheckReader reader) {
    if (g_check_count < kMaxCheckUnits) g_check_units[g_check_count++] = {file, reader};
    return g_check_count;
}
}  // namespace lig

extern "C" int lig_check_status(unsigned long long *failures, int *first_line,
                                const char **first_file) {
#ifdef LIG_CHECKED
    if (!failures || !first_line || !first_file) return LIG_ERR_INVALID_ARGUMENT;
    if (hipDeviceSynchronize() != hipSuccess) return LIG_ERR_LAUNCH;
    *failures = 0;
    *first_line = 0;
    *first_file = nullptr;
    for (int i = 0; i < lig::g_check_count; ++i) {
        unsigned count = 0, line = 0;
        if (lig::g_check_units[i].read(&count, &line, 1) != 0) return LIG_ERR_LAUNCH;
        if (count && !*first_file) {
            *first_file = lig::g_check_units[i].file;
            *first_line = (int)line;
        }
        *failures += count;
    }
    return LIG_OK;
#else
    (void)failures;
    (void)first_line;
    (void)first_file;
    return LIG_ERR_UNSUPPORTED;
#endif
}
