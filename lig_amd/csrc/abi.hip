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
