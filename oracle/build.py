"""Build recipe of the CPU oracle (test infrastructure): oracle/_build/librefcpu.so.

Compiled with g++ -O3 for x86-64-v3 (AVX2/FMA) so that the library built here
also runs on the GPU box's host CPU, where bench.py times it as the CPU
baseline.  The reference itself is not buildable here (Eigen3, FFTW3, Boost,
METIS, NetCDF and gfortran are absent), so there is no oracle/_ref.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(BUILD, "librefcpu.so")


def build():
    os.makedirs(BUILD, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = ["g++", "-O3", "-march=x86-64-v3", "-std=c++17", "-fPIC", "-shared", "-Wall",
           "-I" + os.path.join(ROOT, "include"), os.path.join(HERE, "refcpu.cpp"),
           "-o", tmp, "-lpthread"]
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
