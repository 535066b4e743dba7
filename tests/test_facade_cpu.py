"""include/solidcloth.hpp compiles on its own (with and without the Physics
namespace aliases) and the drape program built from it links against the
library; running it needs a GPU (tests/test_gpu_facade.py)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("defs", [[], ["-DSOLIDCLOTH_NO_PHYSICS_NAMESPACE"]])
def test_facade_header_compiles(defs, tmp_path):
    src = tmp_path / "use.cpp"
    ns = "solidcloth" if defs else "Physics"
    src.write_text(f'#include "solidcloth.hpp"\n'
                   f"void use({ns}::ClothPatch& p) {{ p.GetModel().SetCollisionFlags(0); p.Step(1.f / 60); }}\n")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", *defs,
                        "-I", os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_facade_program_links():
    exe = os.path.join(ROOT, "tests", "cpp", "facade_drape")
    if not os.path.exists(exe):
        pytest.skip("tests/cpp/facade_drape not built (run __graft_entry__.build())")
    r = subprocess.run(["ldd", exe], capture_output=True, text=True)
    assert "libsolidcloth.so" in r.stdout and "not found" not in r.stdout, r.stdout
