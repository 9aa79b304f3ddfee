"""The header-only gamax C++ mirror (eiger_amd/include/gamax/: gamax.hpp,
phys.hpp, pullops.hpp):
compiles and links against the C ABI here; the reference-style C++ test
program (tests/cpp/gamax_shim_test.cpp) runs on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "eiger_amd", "lib")
SRC = os.path.join(ROOT, "tests", "cpp", "gamax_shim_test.cpp")


def _build(out_dir):
    if not os.path.exists(os.path.join(LIBDIR, "libeiger_groupby.so")):
        pytest.fail("libeiger_groupby.so is not built (run __graft_entry__.build())")
    exe = os.path.join(str(out_dir), "gamax_shim_test")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "eiger_amd", "include"), SRC,
           "-L" + LIBDIR, "-leiger_groupby", "-Wl,-rpath," + LIBDIR, "-L/opt/rocm/lib", "-lamdhip64",
           "-Wl,-rpath,/opt/rocm/lib", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    return exe


def test_shim_compiles_and_links(tmp_path):
    exe = _build(tmp_path)
    assert os.path.getsize(exe) > 0


@pytest.mark.gpu
def test_reference_style_cpp_cases_pass_on_gpu(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=600)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "ALL PASSED" in r.stdout
    # the pull-based GroupBy operator's emitted rows against the oracle's
    # group-by of the same input (SUM, COUNT_ALL, MAX over int64 values)
    import numpy as np

    import oracle as O

    def rd(name, dt):
        return np.fromfile(os.path.join(str(tmp_path), name + ".bin"), dtype=dt)

    keys, vals = rd("pullops_in_keys", np.int32), rd("pullops_in_vals", np.int64)
    got = [rd("pullops_out_keys", np.int32), rd("pullops_out_sum", np.int64), rd("pullops_out_count", np.int64),
           rd("pullops_out_max", np.int64)]
    assert len(keys) == 1_000_000 and len(np.unique(got[0])) == len(got[0])
    exp = O.groupby(keys, [vals], [(0, O.SUM), (-1, O.COUNT_ALL), (0, O.MAX)])
    got = O.sort_by_key(got)
    for g, e in zip(got, exp):
        assert g.dtype == e.dtype
        np.testing.assert_array_equal(g, e)
