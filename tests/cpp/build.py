"""Builds the C++ checks under tests/cpp (test infrastructure):

test_host_api      the reference's call sequences on the C++ mirror classes,
                   linked against libmkfbt.so and, as the checker, the FP64
                   oracle (oracle/liboracle.so);
test_adjoint_local the host-only check of the joint-local reverse (hipcc:
                   the header is host- and device-callable);
sanitize_host      the library's host-side sources (pf_io.cpp, pf_model.cpp,
                   pf_host.cpp) and the oracle built with g++/gcc under
                   AddressSanitizer and UndefinedBehaviorSanitizer.

Both need libmkfbt.so and liboracle.so built first."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
PKG = os.path.join(ROOT, "mkfbodytracker_pdaf_amd")
ORACLE = os.path.join(ROOT, "oracle")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HOST_API_SRC = os.path.join(HERE, "test_host_api.cpp")
HOST_API_BIN = os.path.join(HERE, "test_host_api")
SAN_SRC = os.path.join(HERE, "sanitize_host.cpp")
SAN_BIN = os.path.join(HERE, "sanitize_host")
SAN_FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]
ADJ_SRC = os.path.join(HERE, "test_adjoint_local.cpp")
ADJ_BIN = os.path.join(HERE, "test_adjoint_local")


def _run(cmd, what):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"{what} failed:\n" + res.stderr[-4000:])


def build_host_api_test() -> str:
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), "-I", ORACLE, HOST_API_SRC,
          "-L", PKG, "-lmkfbt", "-L", ORACLE, "-loracle",
          "-Wl,-rpath,$ORIGIN/../../mkfbodytracker_pdaf_amd", "-Wl,-rpath,$ORIGIN/../../oracle",
          "-o", HOST_API_BIN], "g++ (test_host_api)")
    return HOST_API_BIN


def build_adjoint_local_test() -> str:
    _run([HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", ADJ_SRC, "-o", ADJ_BIN], "hipcc (test_adjoint_local)")
    return ADJ_BIN


def build_sanitize_test() -> str:
    import tempfile
    csrc = os.path.join(PKG, "csrc")
    with tempfile.TemporaryDirectory(prefix="mkfbt_san_") as tmp:
        objs = []
        for src in ("pf_oracle.c", "pf_pose_ref.c"):
            o = os.path.join(tmp, src + ".o")
            _run(["gcc", "-std=c11", "-fopenmp", *SAN_FLAGS, "-c", os.path.join(ORACLE, src), "-o", o], "gcc " + src)
            objs.append(o)
        _run(["g++", "-std=c++17", "-fopenmp", *SAN_FLAGS, "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include",
              "-I", os.path.join(ROOT, "include"), "-I", ORACLE, SAN_SRC, os.path.join(csrc, "pf_io.cpp"),
              os.path.join(csrc, "pf_model.cpp"), os.path.join(csrc, "pf_host.cpp"),
              os.path.join(ORACLE, "pf_adjoint_ref.cpp"), *objs, "-o", SAN_BIN], "g++ (sanitize_host)")
    return SAN_BIN


def build() -> None:
    build_host_api_test()
    build_adjoint_local_test()
    build_sanitize_test()


if __name__ == "__main__":
    build()
