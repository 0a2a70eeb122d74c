"""Build libpd_gpu1_amd.so in-tree with hipcc for gfx950 (see csrc/Makefile)."""
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(_HERE, "csrc")


def build(jobs: int = 8, force: bool = True) -> str:
    """Compile every HIP translation unit and link the shared library.

    Returns the path of the built library.  ``force`` rebuilds from scratch.
    """
    cmd = ["make", "-C", CSRC, f"-j{jobs}"]
    if force:
        cmd.append("-B")
    subprocess.run(cmd, check=True)
    return os.path.join(_HERE, "libpd_gpu1_amd.so")


if __name__ == "__main__":
    print(build())
