"""Build step for the native library: compiles libebbcfd.so and ebb-solve for gfx950
with the Makefile in ebbcfd_amd/csrc (hipcc), always recompiling."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")


def build(jobs=8, force=True):
    if force:
        subprocess.check_call(["make", "-s", "-C", CSRC, "clean"])
    subprocess.check_call(["make", "-s", "-j%d" % jobs, "-C", CSRC, "all"])
    return os.path.join(HERE, "libebbcfd.so")


if __name__ == "__main__":
    print(build())
