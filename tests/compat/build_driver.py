"""Compile tests/compat/driver.cpp against include/vie_compat.hpp and
libviefft.so (g++, host code only) into tests/compat/driver.bin."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "driver.bin")


def build(out=OUT):
    lib = os.path.join(ROOT, "vie_fft_amd")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "driver.cpp"), "-o", out, "-L", lib, "-lviefft", "-Wl,-rpath," + lib,
           "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.run(cmd, check=True)
    return out


if __name__ == "__main__":
    print(build(sys.argv[1] if len(sys.argv) > 1 else OUT))
