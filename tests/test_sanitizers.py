"""Host-side code under the sanitizers (CPU only, no GPU calls).

tests/sanitize/check_host.cpp drives the library's host entry points (the
GNM generator, the threaded RGG2D / RHG generators, the threaded difference
encoder, the argument checks of the C ABI) and the oracle (Kruskal, both
Boruvka forms, filter-Boruvka at several thresholds, error statuses).  It is
built twice, each time with the library sources compiled for gfx950 with
host-side sanitizer flags only (-Xarch_host; device code is not
instrumented):
  * AddressSanitizer + UndefinedBehaviorSanitizer, leaks checked;
  * ThreadSanitizer, the generators and encoder on their worker threads.
A planted heap overflow checks that the AddressSanitizer build really
reports.  The reference wires the same sanitizers into its CMake build
(CMakeLists.txt:51-62).
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "kamsta_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CLANG = "/opt/rocm/llvm/bin/clang++"
LLVM_LIB = "/opt/rocm/llvm/lib"
CHECK_SRC = os.path.join(HERE, "sanitize", "check_host.cpp")
ORACLE_SRC = os.path.join(ROOT, "oracle", "kmst_oracle.cpp")

SANITIZERS = {"asan": ["address", "undefined"], "tsan": ["thread"]}

pytestmark = pytest.mark.skipif(not (os.path.exists(HIPCC) and os.path.exists(CLANG)),
                                reason="hipcc / ROCm clang++ not installed")


def _library_cmd(kind, out):
    from kamsta_amd import build
    host = []
    for s in SANITIZERS[kind]:
        host += ["-Xarch_host", "-fsanitize=" + s]
    host += ["-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-g"]
    return [HIPCC, "--offload-arch=gfx950", "-O1", "-std=c++17", "-fPIC", "-shared",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", *host, "-o", out,
            *[os.path.join(CSRC, s) for s in build.SOURCES]]


def _check_cmd(kind, lib_dir, out):
    return [CLANG, "-O1", "-g", "-std=c++17", "-fopenmp", "-fsanitize=" + ",".join(SANITIZERS[kind]),
            "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
            "-I", os.path.join(ROOT, "include"), "-o", out, CHECK_SRC, ORACLE_SRC,
            "-L", lib_dir, "-lkmst_" + kind, "-Wl,-rpath," + lib_dir,
            "-L", LLVM_LIB, "-Wl,-rpath," + LLVM_LIB]


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("sanitize"))
    # both library builds at once (about 45 s each)
    procs = {kind: subprocess.Popen(_library_cmd(kind, os.path.join(d, "libkmst_%s.so" % kind)),
                                    stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
             for kind in SANITIZERS}
    for kind, p in procs.items():
        out, _ = p.communicate(timeout=900)
        assert p.returncode == 0, out.decode()[-3000:]
    exes = {}
    for kind in SANITIZERS:
        exe = os.path.join(d, "check_" + kind)
        r = subprocess.run(_check_cmd(kind, d, exe), capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stderr[-3000:]
        exes[kind] = exe
    supp = os.path.join(d, "lsan.supp")
    with open(supp, "w") as f:
        f.write("leak:libomp.so\n")  # OpenMP runtime's own allocations at exit
    return d, exes, supp


def _run(cmd, env_extra, timeout=600):
    env = dict(os.environ)
    env.update(env_extra)
    return subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env=env)


def test_host_code_under_address_and_undefined_sanitizers(built):
    d, exes, supp = built
    r = _run([exes["asan"]], {"ASAN_OPTIONS": "detect_leaks=1",
                              "LSAN_OPTIONS": "suppressions=" + supp,
                              "UBSAN_OPTIONS": "print_stacktrace=1",
                              "OMP_NUM_THREADS": "4"})
    assert r.returncode == 0 and "ok (address+undefined)" in r.stdout, r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr


def test_host_threads_under_thread_sanitizer(built):
    d, exes, _ = built
    # the oracle runs on one OpenMP thread here: the OpenMP runtime itself is
    # not instrumented
    r = _run([exes["tsan"], "threads"], {"TSAN_OPTIONS": "halt_on_error=1 "
                                                         "ignore_noninstrumented_modules=1",
                                         "OMP_NUM_THREADS": "1"})
    assert r.returncode == 0 and "ok (threads)" in r.stdout, r.stderr[-4000:]
    assert "WARNING: ThreadSanitizer" not in r.stderr


def test_address_sanitizer_reports_a_planted_overflow(tmp_path):
    src = tmp_path / "overflow.cpp"
    src.write_text("#include <cstdlib>\n"
                   "int main(int argc, char**) {\n"
                   "  int* p = static_cast<int*>(std::malloc(4 * sizeof(int)));\n"
                   "  p[3 + argc] = 1;  // one past the end\n"
                   "  int r = p[0];\n"
                   "  std::free(p);\n"
                   "  return r;\n"
                   "}\n")
    exe = str(tmp_path / "overflow")
    r = subprocess.run([CLANG, "-O0", "-g", "-fsanitize=address", "-o", exe, str(src)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    r = _run([exe], {"ASAN_OPTIONS": "detect_leaks=0"})
    assert r.returncode != 0 and "heap-buffer-overflow" in r.stderr
