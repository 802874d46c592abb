"""Runs the CPU tests (pytest -m "not gpu") under AddressSanitizer and
UndefinedBehaviorSanitizer: builds the host code of libqed3.so and the oracle instrumented
(qed3_amd/_asan/, oracle/_asan/, both with the ROCm clang so one sanitizer runtime serves
both), then starts pytest with that runtime preloaded and QED3_LIB / QED3_ORACLE_LIB
pointing at the instrumented builds.  Leak detection is off (the interpreter itself keeps
allocations to the end); any other report stops the run with a non-zero status.

Usage: python tests/run_sanitized.py [extra pytest arguments]
"""
import os
import pathlib
import subprocess
import sys

ROOT = pathlib.Path(__file__).resolve().parents[1]
CLANG = "/opt/rocm/llvm/bin/clang++"


def main():
    sys.path.insert(0, str(ROOT))
    from qed3_amd import build as B
    lib = B.build(verbose=True, sanitize=True)
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "asan"], check=True)
    runtime = subprocess.run([CLANG, "-print-file-name=libclang_rt.asan-x86_64.so"], check=True,
                             capture_output=True, text=True).stdout.strip()
    if not os.path.exists(runtime):
        raise SystemExit(f"sanitizer runtime not found: {runtime}")
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": runtime,
        "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=0",
        "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1",
        "QED3_LIB": str(lib),
        "QED3_ORACLE_LIB": str(ROOT / "oracle" / "_asan" / "liboracle.so"),
    })
    cmd = [sys.executable, "-m", "pytest", str(ROOT / "tests"), "-q", "-x", "-m", "not gpu",
           "-p", "no:cacheprovider"] + sys.argv[1:]
    sys.exit(subprocess.run(cmd, env=env).returncode)


if __name__ == "__main__":
    main()
