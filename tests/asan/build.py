"""Build the AddressSanitizer check of the library's host code:

* tests/asan/libviefft_asan.so -- the product library with vie_api.hip (the C
  ABI and every host-side step) compiled with AddressSanitizer on its host
  code only (device code is that of the product library; the kernel
  translation units' objects are reused as built);
* tests/asan/abi_driver.bin -- tests/asan/abi_driver.cpp, a host program that
  calls the ABI end to end, linked with the sanitizer runtime (clang, the
  same compiler as hipcc's host pass).

The product library must be built first (vie_fft_amd/build.py)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from vie_fft_amd import build as B  # noqa: E402

LIB = os.path.join(HERE, "libviefft_asan.so")
DRIVER = os.path.join(HERE, "abi_driver.bin")
CLANG = "/opt/rocm/lib/llvm/bin/clang++"


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd[:3])} ... failed:\n{r.stderr[-4000:]}")


def build() -> str:
    objs = sorted(os.path.join(B.OBJ, f) for f in os.listdir(B.OBJ) if f.endswith(".o")) if os.path.isdir(B.OBJ) else []
    if not any(o.endswith("vie_api.o") for o in objs):
        B.build()
        objs = sorted(os.path.join(B.OBJ, f) for f in os.listdir(B.OBJ) if f.endswith(".o"))
    api = os.path.join(HERE, "vie_api_asan.o")
    # host code only: every sanitizer flag goes to the host pass
    _run([B.HIPCC, *B.FLAGS, "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer",
          "-Xarch_host", "-g", "-c", os.path.join(B.SRC, "vie_api.hip"), "-o", api])
    others = [o for o in objs if not o.endswith("vie_api.o")]
    _run([B.HIPCC, "-shared", "-fPIC", f"--offload-arch={B.ARCH}", api, *others, "-o", LIB])
    os.remove(api)
    _run([CLANG, "-O1", "-g", "-std=c++17", "-fno-omit-frame-pointer", "-fsanitize=address",
          "-I", os.path.join(ROOT, "include"), os.path.join(HERE, "abi_driver.cpp"), "-o", DRIVER,
          "-L", HERE, "-l:libviefft_asan.so", "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-lamdhip64",
          "-Wl,-rpath,/opt/rocm/lib"])
    return DRIVER


if __name__ == "__main__":
    print(build())
