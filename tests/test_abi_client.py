"""include/blqh.h is a C ABI: a C99 client and a C++17 client compile
against it and link against libblqh.so (built, not run: no GPU needed)."""
import os
import shutil
import subprocess

import pytest

import blqh_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "abi_client", "client.c")


@pytest.mark.parametrize("compiler,std,lang", [("gcc", "-std=c99", "c"), ("g++", "-std=c++17", "c++")])
def test_client_compiles_and_links(tmp_path, compiler, std, lang):
    if shutil.which(compiler) is None:
        pytest.skip(f"{compiler} not available")
    lib = blqh_amd.lib()._name
    exe = tmp_path / f"client_{lang}"
    cmd = [compiler, std, "-Wall", "-Wextra", "-Werror", "-x", lang, SRC, "-x", "none", "-I",
           os.path.join(ROOT, "include"), "-o", str(exe), lib, f"-Wl,-rpath,{os.path.dirname(lib)}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert exe.exists()
