"""In-tree build of libkdhip.so (HIP kernels for gfx950 + C ABI + host tree builder)."""
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs=8):
    """Compile every HIP/C++ source of the library (always rebuilds)."""
    csrc = os.path.join(_HERE, "csrc")
    subprocess.run(["make", "-s", "-C", csrc, "clean"], check=True)
    subprocess.run(["make", "-s", "-C", csrc, f"-j{jobs}"], check=True)
    adapter = os.path.join(_HERE, "adapter")
    subprocess.run(["make", "-s", "-C", adapter, "clean"], check=True)
    subprocess.run(["make", "-s", "-C", adapter], check=True)
    return os.path.join(_HERE, "libkdhip.so")
