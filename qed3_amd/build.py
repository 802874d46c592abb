"""Build libqed3.so (HIP for gfx950 + host C++) in-tree with hipcc.

Usage: python -m qed3_amd.build  (or qed3_amd.build.build()).  Objects go to
qed3_amd/_build/, the library to qed3_amd/libqed3.so.
"""
import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent
CSRC = HERE / "csrc"
OBJ = HERE / "_build"
LIB = HERE / "libqed3.so"
ARCH = "gfx950"


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and pathlib.Path(c).exists():
            return c
    raise RuntimeError("hipcc not found")


def _common_flags():
    return [
        "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
        # keep packed FP32 instructions out of the device code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", str(ROOT / "include"), "-Wall", "-Wno-unused-result",
    ]


# AddressSanitizer + UndefinedBehaviorSanitizer on the host code only (the device code is
# built as usual); the library then needs the sanitizer runtime preloaded into the process.
SAN_OBJ = HERE / "_build_asan"
SAN_LIB = HERE / "_asan" / "libqed3.so"
SAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-g"]


def _compile(src: pathlib.Path, sanitize: bool = False) -> pathlib.Path:
    obj = (SAN_OBJ if sanitize else OBJ) / (src.name + ".o")
    cmd = [_hipcc()] + _common_flags() + (SAN_FLAGS if sanitize else []) + ["-c", str(src), "-o", str(obj)]
    if src.suffix == ".cpp":
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False, sanitize: bool = False) -> pathlib.Path:
    """Compile every source and link libqed3.so; sanitize=True builds the host-instrumented
    copy into qed3_amd/_asan/ instead (for the CPU tests, tests/run_sanitized.py)."""
    obj_dir, lib = (SAN_OBJ, SAN_LIB) if sanitize else (OBJ, LIB)
    obj_dir.mkdir(exist_ok=True)
    lib.parent.mkdir(exist_ok=True)
    srcs = sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, sanitize), srcs))
    tmp = lib.with_suffix(".so.tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp)] + (
        SAN_FLAGS if sanitize else []) + [str(o) for o in objs] + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)
    if verbose:
        print(f"built {lib}")
    return lib


if __name__ == "__main__":
    build(verbose=True, sanitize="--sanitize" in sys.argv[1:])
    sys.exit(0)
