"""eiger_amd -- MI355X-native single-key grouped aggregation (hash, sort and
partition group-by plus the radix-partition and atomics sub-operators).

The compute path is libeiger_groupby.so (HIP C++ for gfx950) behind the C ABI in
include/eiger_groupby.h.  ``eiger_amd.gamax`` mirrors the reference's gamax
operator API over that ABI; the header-only C++ mirror lives in
eiger_amd/include/gamax/.
"""
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))


def build(force: bool = True, jobs: int = 8) -> None:
    """Compile libeiger_groupby.so for gfx950 (hipcc, in-tree)."""
    cmd = ["make", "-C", os.path.join(_HERE, "csrc"), f"-j{jobs}"]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True)


from . import gamax  # noqa: E402,F401
