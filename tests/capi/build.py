"""Builds tests/capi/caller.cpp (test infrastructure: a C++ caller of the C
ABI) with g++ against include/ and the in-tree liblmx_mpd.so."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build", "capi_caller")


def build() -> str:
    libdir = os.path.join(ROOT, "lattice_microbes_amd", "_build")
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    subprocess.run([os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
                    os.path.join(HERE, "caller.cpp"), "-L", libdir, "-llmx_mpd",
                    "-Wl,-rpath,$ORIGIN/../../../lattice_microbes_amd/_build",
                    "-o", OUT], check=True)
    return OUT


if __name__ == "__main__":
    print(build())
