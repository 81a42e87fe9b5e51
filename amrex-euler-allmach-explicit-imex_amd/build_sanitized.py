"""CPU-only AddressSanitizer + UndefinedBehaviorSanitizer builds of the
host-side code, for running the CPU test suite under the sanitizers.

Outputs, all under _san/ next to this file (and oracle/_san/):
  libec_amd.so       the product library with its host code instrumented
                     (ec_problems.cpp, ec_adjoint_step_host, the C-ABI entry
                     points' argument checks); device code is compiled as in
                     build.py and is not instrumented
  euler_compare_amd  the drop-in driver (its inputs-file parser and checks)
  ../oracle/_san/libec_oracle.so   the CPU oracle, built with the same clang
                     so that one sanitizer runtime serves the whole process

Every instrumented object links the shared clang runtime
(libclang_rt.asan-x86_64.so), which tests/run_sanitized.sh preloads into the
Python process before pointing the loaders at these files (EC_AMD_LIB,
EC_ORACLE_LIB, EC_AMD_DRIVER).  Float division by zero is not a check of
-fsanitize=undefined: the code relies on IEEE infinities and NaNs there.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

from build import ARCH, CSRC, DEVICE_FLAGS, HIPCC, PKG

SAN = PKG / "_san"
ROOT = PKG.parent
ORACLE = ROOT / "oracle"
CLANG = os.environ.get("EC_SAN_CLANG", "/opt/rocm/llvm/bin/clang")

# each host-only sanitizer flag follows -Xarch_host, so the device code is
# compiled exactly as in build.py
HOST_SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-fno-sanitize-recover=undefined"]
LINK_SAN = [*HOST_SAN[:4], "-shared-libasan"]
CLANG_SAN = ["-fsanitize=address,undefined", "-shared-libasan"]  # plain clang, host code only


def _compile(src: Path) -> Path:
    obj = SAN / "obj" / (src.stem + ".o")
    cmd = [HIPCC, *DEVICE_FLAGS, *HOST_SAN, "-g", "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src.name}")
    return obj


def runtime() -> Path:
    """The shared sanitizer runtime the instrumented files link."""
    out = subprocess.run([CLANG, "-print-file-name=libclang_rt.asan-x86_64.so"],
                         capture_output=True, text=True, check=True).stdout.strip()
    p = Path(out)
    if not p.exists():
        raise RuntimeError(f"sanitizer runtime not found: {out}")
    return p


def build(jobs: int | None = None) -> dict:
    (SAN / "obj").mkdir(parents=True, exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    lib = SAN / "libec_amd.so"
    subprocess.run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *LINK_SAN, "-o", str(lib),
                    *map(str, objs)], check=True)
    exe = SAN / "euler_compare_amd"
    subprocess.run([HIPCC, f"--offload-arch={ARCH}", "-O1", "-g", "-std=c++17", "-ffp-contract=off",
                    *HOST_SAN, "-shared-libasan",
                    str(PKG / "host" / "euler_compare_amd.cpp"), "-o", str(exe), f"-L{SAN}", "-lec_amd",
                    "-Wl,-rpath,$ORIGIN"], check=True)
    osan = ORACLE / "_san"
    osan.mkdir(exist_ok=True)
    olib = osan / "libec_oracle.so"
    subprocess.run([CLANG, "-O1", "-g", "-fPIC", "-std=c11", "-ffp-contract=off", "-fno-fast-math",
                    "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined", *CLANG_SAN, "-shared",
                    "-o", str(olib), str(ORACLE / "ec_oracle.c"), str(ORACLE / "ec_oracle_imex.c"),
                    str(ORACLE / "ec_oracle_direct.c"), "-lm"], check=True)
    return {"lib": lib, "driver": exe, "oracle": olib, "runtime": runtime()}


if __name__ == "__main__":
    for k, v in build().items():
        print(f"{k}={v}")
