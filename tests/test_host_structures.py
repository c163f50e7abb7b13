"""Host-side data structures of the drop-in classes (Rice descriptors, Elias-Fano, retrieval), built
from the package's headers with g++ (no GPU)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_structures(tmp_path):
    exe = tmp_path / "hs"
    subprocess.run(["g++", "-O2", "-std=c++20", "-I", os.path.join(ROOT, "shockhash_amd", "include"),
                    "-o", str(exe), os.path.join(ROOT, "tests", "native", "host_structures.cpp")], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
