"""CPU compile check of the MEX shims (swd_amd/mex/*.cpp, the drop-ins for
the reference's eigen_staggerfd_single/double, eigen_e2drtm_single/double
and the RT*/LHDispPick/ADWDgrad_1 gateways).  MATLAB is not installed, so
they are compiled with -fsyntax-only against tests/mex_stub/, which declares
the MATLAB C++ Data API subset they use: a type or signature error in the
drop-in layer, or a mismatch with include/swd_abi.h, fails here."""
import glob
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIMS = sorted(glob.glob(os.path.join(ROOT, "swd_amd", "mex", "*.cpp")))


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
@pytest.mark.parametrize("src", SHIMS, ids=[os.path.basename(s) for s in SHIMS])
def test_mex_shim_compiles_against_the_data_api(src):
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-Wno-unused-variable",
           "-I" + os.path.join(ROOT, "tests", "mex_stub"), "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "swd_amd", "mex"), src]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_every_shim_is_checked():
    names = {os.path.basename(s) for s in SHIMS}
    for n in ("eigen_staggerfd_single.cpp", "eigen_staggerfd_double.cpp", "eigen_e2drtm_single.cpp",
              "eigen_e2drtm_double.cpp", "swd_rt_mex.cpp", "swd_lhdisppick_mex.cpp", "swd_adwdgrad1_mex.cpp"):
        assert n in names
