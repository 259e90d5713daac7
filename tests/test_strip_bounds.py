"""CPU: every global-memory access of the solver's strip walk stays inside its
plane.  tests/cpp/strip_bounds.cpp enumerates the addresses that
k_irls_weights, k_cg_start, k_cg_iter (both walk directions) and
k_refine_finalize form, with the kernels' own grid mapping, clamped load
index and walk bounds (csrc/strip_geom.h), over strip heights 2 .. 32,
factored and explicit carriers, shapes 1x1 .. 313x1135 alone, three of a
size, mixed, and the largest plane the solver accepts (992 x 2048).  With
the loads formed unclamped (--plant) the same check must report accesses
outside the plane: it sees the class of fault the clamp removes."""
import pathlib
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    exe = tmp_path_factory.mktemp("strip") / "strip_bounds"
    subprocess.run([cxx, "-std=c++17", "-O2", "-Wall", "-Werror", f"-I{ROOT / 'hvd-cvbs-decoding_amd' / 'csrc'}",
                    "-o", str(exe), str(ROOT / "tests" / "cpp" / "strip_bounds.cpp")], check=True)
    return exe


def test_strip_walk_addresses_in_bounds(checker):
    r = subprocess.run([str(checker)], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert " 0 outside their planes" in r.stdout


def test_strip_walk_check_sees_unclamped_loads(checker):
    r = subprocess.run([str(checker), "--plant"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 1, r.stdout
    assert "walk load" in r.stdout
