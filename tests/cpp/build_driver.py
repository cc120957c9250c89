"""Builds tests/cpp/_build/mgsol_driver (the C++ host-layer driver) against the in-tree libfemus_hip.so, and
the host checkers tests/cpp/_build/sell_layout_check (SELL layout), tests/cpp/_build/order_check (internal
DoF order), tests/cpp/_build/maps_dump (DoF maps, patterns and transfers of the mesh unit) and
tests/cpp/_build/egather_check (the element-buffer gather plan and kernel)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build", "mgsol_driver")


CHECK = os.path.join(HERE, "_build", "sell_layout_check")
ORDER = os.path.join(HERE, "_build", "order_check")
MAPS = os.path.join(HERE, "_build", "maps_dump")
EGATHER = os.path.join(HERE, "_build", "egather_check")
CSRC = os.path.join(ROOT, "exfemus_amd", "csrc")


def build_check():
    """The host checkers, each built with the library's own host units (no GPU): sell_layout_check.cpp with the
    SELL layout builder, order_check.cpp with the mesh builder and the internal DoF order, maps_dump.cpp with the
    mesh builder's DoF maps, patterns and transfers, egather_check.cpp with the gather planner, the mesh unit and
    the SELL builder."""
    os.makedirs(os.path.dirname(CHECK), exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", "-fopenmp", "-Wall", os.path.join(HERE, "sell_layout_check.cpp"),
           os.path.join(CSRC, "sell_host.cpp"), "-o", CHECK]
    subprocess.run(cmd, check=True)
    cmd = ["g++", "-O2", "-std=c++17", "-fopenmp", "-Wall", os.path.join(HERE, "order_check.cpp"),
           os.path.join(CSRC, "gencase.cpp"), os.path.join(CSRC, "order_host.cpp"), "-o", ORDER]
    subprocess.run(cmd, check=True)
    cmd = ["g++", "-O2", "-std=c++17", "-fopenmp", "-Wall", os.path.join(HERE, "maps_dump.cpp"),
           os.path.join(CSRC, "gencase.cpp"), "-o", MAPS]
    subprocess.run(cmd, check=True)
    cmd = ["g++", "-O2", "-std=c++17", "-fopenmp", "-Wall", os.path.join(HERE, "egather_check.cpp"),
           os.path.join(CSRC, "egather_host.cpp"), os.path.join(CSRC, "gencase.cpp"),
           os.path.join(CSRC, "sell_host.cpp"), "-o", EGATHER]
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
