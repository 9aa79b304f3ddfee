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
