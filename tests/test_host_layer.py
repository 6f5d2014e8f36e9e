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
                    "-L", libdir, "-lg5", "-lg5h5", "-Wl,-rpath," + libdir], check=True)
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


def test_ini_jet_inlet(exe):
    # [border] enable_jet: the jet inlet with the region-1 state and the [jet] geometry (defaults
    # of border_conditions/BCJet.h:62-69 for the keys the file leaves out)
    out = subprocess.run([exe, os.path.join(GOLD, "jet_fuel_n2_2d_fixed.ini")], check=True,
                         capture_output=True, text=True).stdout
    kv = parse(out)
    assert kv["inlet"] == "2" and kv["nb"] == "12,4,1"
    y, z, w = (float(x) for x in kv["jet"].split(","))
    assert (y, z, w) == (0.5, 0.5, 0.25)
    u = [float(x) for x in kv["inlet_u"].split(",")]
    # region 1: alpha0 = 0, alpha_rho = (0, 623), p = 3.83982e5, u = 18 (fuel: gamma 4.4, pinf 8e7)
    rho_e = (3.83982e5 + 4.4 * 8e7) / 3.4
    assert u[0] == 623.0 and u[2] == 623.0 * 18.0 and u[3] == 0.0
    assert u[4:] == [0.0, 0.0, 1.0, 623.0]
    assert abs(u[1] - (rho_e + 0.5 * 623.0 * 18.0 ** 2)) <= 1e-12 * u[1]
    # without the key no inlet is set
    out = subprocess.run([exe, os.path.join(GOLD, "shock_bubble_air_helium_2d_512x128.ini")], check=True,
                         capture_output=True, text=True).stdout
    assert parse(out)["inlet"] == "0"


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
def test_cpp_checkpoint_restart(exe, tmp_path):
    # save_solution_hdf5 after 30 iterations, init_restart of a second solver from that file, 20
    # more iterations on each: identical states, times and iteration counts
    print("[gpu] C++ save_solution_hdf5 + init_restart on the 2-D fixture", flush=True)
    out = subprocess.run([exe, os.path.join(GOLD, "shock_bubble_air_helium_2d_512x128.ini"), "--restart", str(tmp_path)],
                         capture_output=True, text=True, timeout=120)
    sys.stdout.write(out.stdout + out.stderr)
    assert out.returncode == 0
    kv = parse(out.stdout.splitlines()[-1])
    assert kv["identical"] == "1" and kv["restart_iteration"] == "30" and kv["iteration"] == "50"
    assert os.path.exists(os.path.join(tmp_path, "ckpt_0000030.xmf"))


@pytest.mark.gpu
def test_cpp_pool_class_runs(exe):
    print("[gpu] C++ PoolHip over seeded batch solves", flush=True)
    out = subprocess.run([exe, "--pool", "6"], capture_output=True, text=True, timeout=120)
    sys.stdout.write(out.stdout + out.stderr)
    assert out.returncode == 0
    kv = parse(out.stdout.splitlines()[-1])
    assert kv["complete"] == "6" and kv["finished"] == "6"
