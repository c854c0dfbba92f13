"""Config front-end and host facade on the CPU."""
import os
import subprocess

import pytest

from cbns3d_cart_mpi_amd import parser_ini

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cbns3d_cart_mpi_amd", "bin", "cbns3d_amd")

# the Taylor-Green configuration's sections and keys, in the order
# read_input consumes them (src/read_input.cpp:24-72)
INI = """[string]
checkpoint_fname = checkpoint

[real]
CFL = 0.5
t_cur = 0.0
t_end = 20.0

Pr = 0.71
gamma = 1.4

Mach = 0.1
Re = 1600.0
T_inf = 255.556

L = 1.0

[integer]
IM = 16
JM = 16
KM = 16

nstep_max = 10
checkpoint_freq = 10
num_output = 2

[bool]
is_restart = 1
"""

ORDER = ["checkpoint", "0.5", "0.0", "20.0", "0.71", "1.4", "0.1", "1600.0", "255.556", "1.0",
         "16", "16", "16", "10", "10", "2", "1"]


def test_ini_to_input_txt_positional_order(tmp_path):
    ini = tmp_path / "tgv.ini"
    ini.write_text(INI)
    out = tmp_path / "input.txt"
    vals = parser_ini.convert(str(ini), str(out), echo=False)
    assert vals == ORDER
    assert out.read_text().splitlines() == ORDER


def test_missing_ini_raises(tmp_path):
    with pytest.raises(FileNotFoundError):
        parser_ini.convert(str(tmp_path / "nope.ini"), str(tmp_path / "x.txt"), echo=False)


def test_driver_reports_missing_input(tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("driver not built")
    r = subprocess.run([EXE, "1", "1", "1"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 1
    assert "Exception opening/reading file" in r.stdout


def test_driver_usage_message(tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("driver not built")
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stdout


def test_facade_compiles_the_reference_drivers_block3d_calls():
    """The reference driver's Block3d calls, with its rank-suffixed checkpoint
    name and output_vtr(fname), compile against the facade (compile only)."""
    host = os.path.join(ROOT, "cbns3d_cart_mpi_amd", "host")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                        "-I" + os.path.join(ROOT, "include"), "-I" + host,
                        os.path.join(ROOT, "tests", "facade_api_check.cpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
