"""Host check of the fast cell-index division (pica_amd/csrc/pk_fastdiv.h) against the C++
integer division, by compiling tests/cxx/fastdiv_check.cpp with g++."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_fastdiv_matches_integer_division(tmp_path):
    exe = str(tmp_path / "fastdiv_check")
    subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(HERE, "cxx", "fastdiv_check.cpp"), "-o", exe],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
