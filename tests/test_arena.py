"""CPU test of the serving arena's offset allocator (trident_amd/csrc/arena.h):
tests/cpp/test_arena.cpp is compiled with g++ and checks it against a
brute-force occupancy map."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_arena_allocator(tmp_path):
    exe = str(tmp_path / "test_arena")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", os.path.join(ROOT, "tests", "cpp", "test_arena.cpp"),
                    "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "arena ok" in r.stdout
