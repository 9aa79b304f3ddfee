"""The run-time specialisation compiles without a device: the embedded kernel
sources build through hipRTC for gfx950 (hash and partition kernels, 4- and
8-byte keys, vector and typed-load layouts), and the controls round-trip."""
from eiger_amd import gamax as G


def test_embedded_sources_compile_for_gfx950():
    G.jit_selftest("gfx950")


def test_configure_round_trip():
    before = G.jit_stats()
    try:
        G.jit_configure(G.jit_mode.OFF, 123)
        st = G.jit_stats()
        assert st["mode"] == G.jit_mode.OFF and st["min_rows"] == 123
        G.jit_configure(None, None)
        st = G.jit_stats()
        assert st["mode"] == G.jit_mode.OFF and st["min_rows"] == 123
    finally:
        G.jit_configure(before["mode"], before["min_rows"])
    assert G.jit_stats()["mode"] == before["mode"]


# ---- exit with compiles queued or running ------------------------------------
# A process may end while the background compiler thread is inside hipRTC.
# The library's exit handler runs before the teardown of anything that compile
# uses (it is armed again after every module load and before every compile),
# lets the running compile finish, drops the queue and joins the thread; the
# Python binding also drains the queue in an interpreter exit hook.
import os
import subprocess
import sys
import textwrap

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_REQS = "[(-1, 5), (0, 0), (0, 2), (0, 1), (1, 0)]"  # COUNT, SUM/MIN/MAX(int64), SUM(float32)


def _run(code: str, timeout=240):
    # AMD_COMGR_CACHE=0: the compiler's on-disk cache would make the compiles
    # finish at once, and the exit would not race them
    env = dict(os.environ, AMD_COMGR_CACHE="0")
    return subprocess.run([sys.executable, "-c", textwrap.dedent(code)], cwd=ROOT, capture_output=True, text=True,
                          timeout=timeout, env=env)


def test_prefetch_queues_each_specialisation_once():
    G.jit_wait()
    before = G.jit_stats()
    reqs = [(-1, G.aggregation_id.COUNT_ALL), (0, G.aggregation_id.SUM)]
    n1 = G.jit_prefetch(G.type_id.INT64, [G.type_id.INT32], reqs, arch="gfx950")
    n2 = G.jit_prefetch(G.type_id.INT64, [G.type_id.INT32], reqs, arch="gfx950")
    assert (n1, n2) == (2, 0)  # hash + partition kernels, then already known
    G.jit_wait()
    after = G.jit_stats()
    assert after["compiled"] - before["compiled"] == 2 and after["failed"] == before["failed"]
    assert after["pending"] == 0


def test_exit_while_compiling_without_binding_hook():
    """Raw ctypes (no interpreter exit hook): the process exits 0 while the
    compiles it queued are running or still queued."""
    for delay in (0.0, 0.3):
        r = _run(f"""
            import ctypes as C, time
            L = C.CDLL("eiger_amd/lib/libeiger_groupby.so")
            class Req(C.Structure):
                _fields_ = [("column_index", C.c_int64), ("agg_kind", C.c_int32), ("reserved", C.c_int32)]
            reqs = (Req * 5)(*[(c, k, 0) for c, k in {_REQS}])
            vt = (C.c_int32 * 2)(4, 10)
            q = C.c_int32(0)
            L.eg_jit_prefetch.argtypes = [C.c_int32, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_char_p,
                                          C.c_void_p]
            L.eg_jit_stats.argtypes = [C.c_void_p]
            assert L.eg_jit_prefetch(3, vt, 2, reqs, 5, b"gfx950", C.byref(q)) == 0 and q.value == 2
            time.sleep({delay})
            st = (C.c_int64 * 6)()
            L.eg_jit_stats(st)
            print("pending at exit", st[2], flush=True)
        """)
        assert r.returncode == 0, (delay, r.returncode, r.stderr[-2000:])
        assert "pending at exit" in r.stdout
        assert int(r.stdout.split("pending at exit")[1].split()[0]) >= 1  # exit really raced a compile


def test_binding_drains_compiles_before_native_exit():
    """Through eiger_amd: an interpreter exit hook registered before the
    library is loaded runs after the binding's drain hook (hooks run last in,
    first out) and sees nothing pending."""
    r = _run(f"""
        import atexit
        def check():
            from eiger_amd import gamax as G
            print("pending after drain", G.jit_stats()["pending"], flush=True)
        atexit.register(check)
        from eiger_amd import gamax as G
        assert G.jit_prefetch(G.type_id.INT32, [G.type_id.INT64, G.type_id.FLOAT32], {_REQS}, arch="gfx950") == 2
    """)
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])
    assert "pending after drain 0" in r.stdout
