"""Sanitizer builds of the CPU code (CPU only).

- The OpenMP oracle (oracle/g5_oracle.cpp, driven by tests/cpp/oracle_driver.cpp over 3-D THINC
  with walls, 2-D RK2_SSP with MC slopes, a periodic direction, MILLER_PUCKETT and the shock-bubble
  inlet, 2-D RK1 THINC with the jet inlet) under AddressSanitizer + UndefinedBehaviorSanitizer (GCC,
  leak detection on) and under ThreadSanitizer (LLVM's OpenMP runtime with its archer tool, which
  makes the runtime's own synchronisation visible; its non-instrumented internals are ignored).
- The library's host C++ -- the .ini layer of include/godunov_five_eq.hpp with its refusals, the
  host initial conditions and Rankine-Hugoniot states (g5_init.cpp), the workload specs
  (g5_specs.cpp) and the HDF5 checkpoint round trip (g5_hdf5.cpp) -- driven by
  tests/cpp/host_cpu.cpp under AddressSanitizer + UndefinedBehaviorSanitizer.
Each sanitized run must end cleanly and print the same results as the plain build.
"""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OUT = os.path.join(ROOT, "godunov_five_eq_amd", "_build", "sanitize")
CSRC = os.path.join(ROOT, "godunov_five_eq_amd", "csrc")
ASAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]
COMMON = ["-std=c++17", "-O1", "-g", "-fopenmp", "-ffp-contract=off"]
HDF5 = os.environ.get("G5_HDF5_PREFIX", "/opt/conda")
LLVM = "/opt/rocm/llvm"


def _build(args):
    r = subprocess.run(args, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


@pytest.fixture(scope="module")
def binaries():
    os.makedirs(OUT, exist_ok=True)
    od = os.path.join(ROOT, "tests", "cpp", "oracle_driver.cpp")
    hc = [os.path.join(ROOT, "tests", "cpp", "host_cpu.cpp")] + [os.path.join(CSRC, f) for f in
                                                                   ("g5_init.cpp", "g5_specs.cpp", "g5_hdf5.cpp")]
    libdir = os.path.join(ROOT, "godunov_five_eq_amd")
    # the system library directory first, so that the C++ runtime is not HDF5's older one
    hlink = ["-I", os.path.join(HDF5, "include"), "-L", libdir, "-lg5", os.path.join(HDF5, "lib", "libhdf5.so"),
             f"-Wl,-rpath,/usr/lib/x86_64-linux-gnu:{libdir}:{os.path.join(HDF5, 'lib')}"]
    jobs = {
        "od_plain": ["g++", *COMMON, od, "-o", os.path.join(OUT, "od_plain")],
        "od_asan": ["g++", *COMMON, *ASAN, od, "-o", os.path.join(OUT, "od_asan")],
        "hc_plain": ["g++", *COMMON, *hc, "-o", os.path.join(OUT, "hc_plain"), *hlink],
        "hc_asan": ["g++", *COMMON, *ASAN, *hc, "-o", os.path.join(OUT, "hc_asan"), *hlink],
    }
    tsan = os.path.exists(os.path.join(LLVM, "lib", "libarcher.so"))
    if tsan:
        jobs["od_tsan"] = [os.path.join(LLVM, "bin", "clang++"), *COMMON, "-fsanitize=thread", od, "-o",
                           os.path.join(OUT, "od_tsan"), f"-Wl,-rpath,{os.path.join(LLVM, 'lib')}"]
    with ThreadPoolExecutor(max_workers=5) as ex:
        list(ex.map(_build, jobs.values()))
    return {k: os.path.join(OUT, k) for k in jobs}


def _run(exe, args=(), env=None):
    e = dict(os.environ)
    e.update(env or {})
    e.setdefault("OMP_NUM_THREADS", "4")
    r = subprocess.run([exe, *args], capture_output=True, text=True, env=e, timeout=300)
    return r.returncode, r.stdout, r.stderr


def test_oracle_asan_ubsan(binaries):
    rc, out, err = _run(binaries["od_asan"], env={"ASAN_OPTIONS": "detect_leaks=1"})
    assert rc == 0, err[-3000:]
    rc0, out0, _ = _run(binaries["od_plain"])
    assert rc0 == 0 and out == out0 and len(out.splitlines()) == 3


def test_oracle_tsan_openmp(binaries):
    if "od_tsan" not in binaries:
        pytest.skip("LLVM's OpenMP tool libarcher.so is not installed")
    rc, out, err = _run(binaries["od_tsan"], env={
        "OMP_TOOL_LIBRARIES": os.path.join(LLVM, "lib", "libarcher.so"),
        "TSAN_OPTIONS": "ignore_noninstrumented_modules=1 halt_on_error=1 exitcode=66"})
    assert rc == 0 and "ThreadSanitizer" not in err, err[-3000:]
    _, out0, _ = _run(binaries["od_plain"])
    assert out == out0


def test_host_cpp_asan_ubsan(binaries, tmp_path):
    gold = os.path.join(ROOT, "tests", "golden")
    (tmp_path / "a").mkdir()
    (tmp_path / "p").mkdir()
    rc, out, err = _run(binaries["hc_asan"], [gold, str(tmp_path / "a")], env={"ASAN_OPTIONS": "detect_leaks=1"})
    assert rc == 0, err[-3000:]
    rc0, out0, _ = _run(binaries["hc_plain"], [gold, str(tmp_path / "p")])
    assert rc0 == 0 and out == out0
    assert out.count("refused ") == 6 and "h5_round_trip=1" in out
