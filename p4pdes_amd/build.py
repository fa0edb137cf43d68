"""Build the gfx950 library, the fish driver and the CPU oracle in-tree.

    python -m p4pdes_amd.build          # everything
Outputs: p4pdes_amd/lib/libp4g.so (HIP, C ABI of include/p4g.h),
p4pdes_amd/lib/libp4g_fish.so + p4pdes_amd/lib/fish (driver),
oracle/_build/ (CPU oracle, test infrastructure).
"""
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "p4pdes_amd")
LIB = os.path.join(PKG, "lib")
INC = os.path.join(ROOT, "include")
CSRC = os.path.join(PKG, "csrc")
FISH = os.path.join(PKG, "fish")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("P4G_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are kept out of the device code.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-Wno-unused-value", "-I" + INC]
C_FLAGS = ["-O2", "-std=gnu11", "-fPIC", "-Wall", "-I" + INC, "-I" + FISH]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_lib():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libp4g.so")
    tmp = out + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-shared", "-o", tmp,
                                os.path.join(CSRC, "p4g_solver.hip"),
                                os.path.join(CSRC, "p4g_shim.cpp")])
    os.replace(tmp, out)
    return out


def build_fish():
    os.makedirs(LIB, exist_ok=True)
    srcs = [os.path.join(FISH, f) for f in ("poissonfunctions.c", "problems.c", "fish.c")]
    out = os.path.join(LIB, "libp4g_fish.so")
    _run(["gcc"] + C_FLAGS + ["-shared", "-o", out + ".tmp"] + srcs +
         ["-L" + LIB, "-lp4g", "-Wl,-rpath,$ORIGIN", "-lm"])
    os.replace(out + ".tmp", out)
    exe = os.path.join(LIB, "fish")
    _run(["gcc"] + C_FLAGS + ["-o", exe + ".tmp", os.path.join(FISH, "fish_main.c"),
                              "-L" + LIB, "-lp4g_fish", "-lp4g", "-Wl,-rpath,$ORIGIN", "-lm"])
    os.replace(exe + ".tmp", exe)
    return out


def build_obstacle():
    """libp4g_obstacle.so + the `obstacle` driver (c/ch12/obstacle.c on the ABI),
    linked with the ch6 callbacks as the reference links ../ch6/poissonfunctions.o."""
    os.makedirs(LIB, exist_ok=True)
    obs = os.path.join(PKG, "obstacle")
    srcs = [os.path.join(FISH, "poissonfunctions.c"), os.path.join(obs, "obstacle.c")]
    out = os.path.join(LIB, "libp4g_obstacle.so")
    _run(["gcc"] + C_FLAGS + ["-I" + FISH, "-shared", "-o", out + ".tmp"] + srcs +
         ["-L" + LIB, "-lp4g", "-Wl,-rpath,$ORIGIN", "-lm"])
    os.replace(out + ".tmp", out)
    exe = os.path.join(LIB, "obstacle")
    _run(["gcc"] + C_FLAGS + ["-o", exe + ".tmp", os.path.join(obs, "obstacle_main.c"),
                              "-L" + LIB, "-lp4g_obstacle", "-lp4g", "-Wl,-rpath,$ORIGIN", "-lm"])
    os.replace(exe + ".tmp", exe)
    return out


def build_minimal():
    """libp4g_minimal.so + the `minimal` driver (c/ch7/minimal.c on the ABI)."""
    os.makedirs(LIB, exist_ok=True)
    mdir = os.path.join(PKG, "minimal")
    out = os.path.join(LIB, "libp4g_minimal.so")
    _run(["gcc"] + C_FLAGS + ["-I" + FISH, "-shared", "-o", out + ".tmp", os.path.join(FISH, "poissonfunctions.c"),
                              os.path.join(mdir, "minimal.c"),
                              "-L" + LIB, "-lp4g", "-Wl,-rpath,$ORIGIN", "-lm"])
    os.replace(out + ".tmp", out)
    exe = os.path.join(LIB, "minimal")
    _run(["gcc"] + C_FLAGS + ["-o", exe + ".tmp", os.path.join(mdir, "minimal_main.c"),
                              "-L" + LIB, "-lp4g_minimal", "-lp4g", "-Wl,-rpath,$ORIGIN", "-lm"])
    os.replace(exe + ".tmp", exe)
    return out


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_all():
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found: the gfx950 library cannot be built")
    build_lib()
    build_fish()
    build_obstacle()
    build_minimal()
    build_oracle()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
