"""Builds tests/cpp/_build/mgsol_driver (the C++ host-layer driver) against the in-tree libfemus_hip.so, and
tests/cpp/_build/sell_layout_check (the host SELL layout checker)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build", "mgsol_driver")


CHECK = os.path.join(HERE, "_build", "sell_layout_check")


def build_check():
    """The host SELL layout checker: sell_layout_check.cpp with the library's host layout builder (no GPU)."""
    os.makedirs(os.path.dirname(CHECK), exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", "-fopenmp", "-Wall", os.path.join(HERE, "sell_layout_check.cpp"),
           os.path.join(ROOT, "exfemus_amd", "csrc", "sell_host.cpp"), "-o", CHECK]
    subprocess.run(cmd, check=True)
    return CHECK


def build():
    build_check()
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    lib_dir = os.path.join(ROOT, "exfemus_amd")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "mgsol_driver.cpp"), "-L", lib_dir, "-lfemus_hip",
           "-Wl,-rpath,$ORIGIN/../../../exfemus_amd", "-o", OUT]
    subprocess.run(cmd, check=True)
    return OUT


if __name__ == "__main__":
    print(build())
