"""Writes abi_driver.cpp: a host program that calls every entry point that
include/lig_abi.h declares, so that a build of the library with host
AddressSanitizer and UndefinedBehaviorSanitizer checks the host side of the
C-ABI (argument checks, workspace-size arithmetic, workspace carving, host
arrays).  Argument sets per entry point: all zero / NULL; sizes 3 with NULL
pointers; negative sizes; and, only when no GPU is visible (the calls then
fail at their first HIP call, after the host work), sizes 3 with every device
pointer a placeholder address, workspaces declared 1 GiB and the host arrays
(strides, SSIM window and weights) exactly sized.  Every int-returning call
must return LIG_OK or a LIG_ERR_* code (except the version query and the
backward-shares setting, whose values are not status codes).

    python tests/sanitize/gen_abi_driver.py > tests/sanitize/build/abi_driver.cpp
"""
import os
import re
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
# host arrays, by parameter name (function-qualified where the name is generic),
# sized exactly for the "fake" argument set (sizes 3, win_size 7, levels 2)
HOST_ARRAYS = {"x_strides": ("long long", 4, "1"), "y_strides": ("long long", 4, "1"),
               "win": ("float", 7, "0.1f"), "weights": ("float", 2, "0.5f"),
               "lig_adam_bias_corrections.out": ("float", 6, "0.0f")}


# int-returning functions whose value is not a status code
NOT_STATUS = {"lig_abi_version", "lig_set_backward_shares"}


def declarations():
    text = open(os.path.join(ROOT, "include", "lig_abi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = []
    for ret, name, params in re.findall(r"([A-Za-z_][\w\s\*]*?)\b(lig_\w+)\s*\(([^;{]*?)\)\s*;",
                                        text, flags=re.S):
        ps = []
        for p in params.split(","):
            p = " ".join(p.split())
            if not p or p == "void":
                continue
            m = re.match(r"(.*?)(\w+)$", p)
            ps.append((m.group(1).strip(), m.group(2)))
        out.append((" ".join(ret.split()), name, ps))
    return out


def value(ty, name, mode, func=""):
    if f"{func}.{name}" in HOST_ARRAYS:
        name = f"{func}.{name}"
    if name == "stream":
        return "nullptr"   # the default stream (a placeholder would be dereferenced)
    if "*" in ty:
        if mode == "fake":
            if name in HOST_ARRAYS:
                return "host_" + name.replace(".", "_")
            return "(void *)kFake" if ty.startswith("void") else f"({ty})kFake"
        return "nullptr"
    if ty == "size_t":
        return {"zero": "0", "small": "4096", "neg": "0", "fake": "(size_t)1 << 30"}[mode]
    if ty == "float":
        return {"zero": "0.0f", "small": "1.0f", "neg": "-1.0f", "fake": "1.0f"}[mode]
    # int, int32_t, long long: sizes, flags, counts
    if mode == "fake" and name in ("win_size",):
        return "7"
    if mode == "fake" and name in ("levels",):
        return "2"
    if mode == "fake" and name in ("cholesky", "shares", "batch_depth"):
        return "0"
    return {"zero": "0", "small": "3", "neg": "-1", "fake": "3"}[mode]


def main():
    decls = declarations()
    w = sys.stdout.write
    w('// Generated by tests/sanitize/gen_abi_driver.py; do not edit.\n')
    w('#include <hip/hip_runtime.h>\n#include <stdio.h>\n#include "lig_abi.h"\n\n')
    w("static const uintptr_t kFake = 0x100000;\nstatic int bad = 0;\n")
    for name, (ty, count, init) in HOST_ARRAYS.items():
        w(f"static {ty} host_{name.replace('.', '_')}[{count}] = {{{', '.join([init] * count)}}};\n")
    w("static void check(const char *name, const char *mode, int rc) {\n"
      "    if (rc > 0 || rc < -5) { printf(\"%s (%s): status %d\\n\", name, mode, rc); bad = 1; }\n"
      "}\n\nint main() {\n"
      "    int ndev = 0;\n"
      "    if (hipGetDeviceCount(&ndev) != hipSuccess) ndev = 0;\n"
      "    for (int s = -7; s <= 1; ++s) (void)lig_strerror(s);\n")
    modes = ("zero", "small", "neg", "fake")
    for ret, name, ps in decls:   # size queries and other host-only functions first
        if name != "lig_strerror" and (ret != "int" or name in NOT_STATUS):
            for mode in modes:
                w(f"    (void){name}({', '.join(value(t, n, mode, name) for t, n in ps)});\n")
    w("    if (ndev > 0) {  // placeholder device pointers must never reach a GPU\n"
      "        printf(\"ok (a GPU is visible: only host-only functions called)\\n\");\n"
      "        return 0;\n    }\n")
    for ret, name, ps in decls:
        if ret != "int" or name in NOT_STATUS:
            continue
        for mode in modes:
            args = ", ".join(value(t, n, mode, name) for t, n in ps)
            w(f"    check(\"{name}\", \"{mode}\", {name}({args}));\n")
    w('    printf(bad ? "FAILED\\n" : "ok (%d entry points)\\n", ' + str(len(decls)) + ');\n')
    w("    return bad;\n}\n")


if __name__ == "__main__":
    main()
