"""The library's host-side sources (the yml readers of pf_io.cpp, the model
validation, selection basis and model image of pf_model.cpp, the host-only
members of the C++ mirror classes in pf_host.cpp) and the FP64 oracle, built
with g++/gcc under AddressSanitizer and UndefinedBehaviorSanitizer
(tests/cpp/sanitize_host.cpp, built by tests/cpp/build.py) and driven with
the shipped model files, malformed files and edge-case inputs: the run must
finish with no sanitizer report and every check passing."""
import importlib.util
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "sanitize_host")


def _ensure_built():
    if os.path.exists(BIN):
        return
    spec = importlib.util.spec_from_file_location("_cpp_tests_build", os.path.join(HERE, "cpp", "build.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    mod.build_sanitize_test()


def test_host_code_under_asan_ubsan(tmp_path):
    _ensure_built()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([BIN, os.path.join(HERE, "golden"), str(tmp_path)], capture_output=True, text=True,
                       timeout=300, env=env)
    print(r.stdout[-3000:], r.stderr[-3000:])
    assert r.returncode == 0 and r.stdout.strip().endswith("OK")
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
