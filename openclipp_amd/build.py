"""Builds libOpenCLIPP.so in-tree (hipcc, gfx950) via openclipp_amd/csrc/Makefile."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libOpenCLIPP.so")


def build(jobs=8, force=False):
    if force:
        subprocess.check_call(["make", "-s", "-C", CSRC, "clean"])
    subprocess.check_call(["make", "-s", "-j", str(jobs), "-C", CSRC])
    return LIB


if __name__ == "__main__":
    print(build())
