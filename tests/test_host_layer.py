"""The C++ host layer (include/godunov_five_eq.hpp): .ini -> g5_config mapping with the
reference's keys and defaults, and (on a GPU) the SolverGodunovFiveEq loop over the C ABI."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
EXE = os.path.join(ROOT, "godunov_five_eq_amd", "_build", "host_layer")


@pytest.fixture(scope="module")
def exe():
    os.makedirs(os.path.dirname(EXE), exist_ok=True)
    libdir = os.path.join(ROOT, "godunov_five_eq_amd")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-o", EXE, os.path.join(ROOT, "tests", "cpp", "host_layer.cpp"),
                    "-L", libdir, "-lg5", "-Wl,-rpath," + libdir], check=True)
    return EXE


def parse(out):
    kv = {}
    for tok in out.split():
        if "=" in tok:
            k, v = tok.split("=", 1)
            kv[k] = v
    return kv


def test_ini_2d_config(exe):
    out = subprocess.run([exe, os.path.join(GOLD, "shock_bubble_air_helium_2d_512x128.ini")], check=True,
                         capture_output=True, text=True).stdout
    kv = parse(out)
    assert kv["dim"] == "2" and kv["nb"] == "32,8,1" and kv["bc"].startswith("2,2,3,3")
    assert float(kv["dx"]) == 1.0 / 128
    assert float(kv["cfl"]) == 0.8 and kv["thinc"] == "0" and kv["integrator"] == "0"
    assert float(kv["smallp"]) == 1e-20  # smallp = smallc^2, models/HydroSettings.cpp:24
    assert kv["gamma"] == "1.3999999999999999,1.6479999999999999"


def test_ini_3d_config(exe):
    out = subprocess.run([exe, os.path.join(GOLD, "shock_bubble_air_water_3d_256.ini")], check=True,
                         capture_output=True, text=True).stdout
    kv = parse(out)
    assert kv["dim"] == "3" and kv["nb"] == "16,16,16" and kv["thinc"] == "1"
    assert float(kv["beta"]) == 1.6 and float(kv["dx"]) == 1.0 / 256
    assert kv["pinf"] == "0,600000000"


@pytest.mark.gpu
def test_cpp_solver_loop_runs(exe):
    print("[gpu] C++ SolverGodunovFiveEq on the 2-D fixture", flush=True)
    out = subprocess.run([exe, os.path.join(GOLD, "shock_bubble_air_helium_2d_512x128.ini"), "--run", "200"],
                         capture_output=True, text=True, timeout=120)
    sys.stdout.write(out.stdout + out.stderr)
    assert out.returncode == 0
    kv = parse(out.stdout.splitlines()[-1])
    assert kv["steps"] == "200"
    assert float(kv["Mcell-updates/s"]) > 0


@pytest.mark.gpu
def test_cpp_pool_class_runs(exe):
    print("[gpu] C++ PoolHip over seeded batch solves", flush=True)
    out = subprocess.run([exe, "--pool", "6"], capture_output=True, text=True, timeout=120)
    sys.stdout.write(out.stdout + out.stderr)
    assert out.returncode == 0
    kv = parse(out.stdout.splitlines()[-1])
    assert kv["complete"] == "6" and kv["finished"] == "6"
