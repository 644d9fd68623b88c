"""Regenerates mt19937_64_libstdcxx.json from mt_golden.cpp (g++, C++ <random>)."""
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))

if __name__ == "__main__":
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "mt_golden")
        subprocess.check_call(["g++", "-O2", "-std=c++11", "-o", exe, os.path.join(HERE, "mt_golden.cpp")])
        out = subprocess.check_output([exe]).decode()
    with open(os.path.join(HERE, "mt19937_64_libstdcxx.json"), "w") as f:
        f.write(out)
    print("wrote", len(out), "bytes")
