"""MI355X-native agent-based simulation core (FLAME GPU 2 compatible API).

The compute path is the C-ABI library libfgx.so (hand-written HIP kernels for
gfx950) and the model libraries built on the C++ API in flamegpu2_amd/include.
This package only locates and loads them; it has no CPU fallback.
"""
import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(PKG, "lib")


class NativeLibraryMissing(RuntimeError):
    pass


def lib_path(name):
    return os.path.join(LIBDIR, name)


def load(name):
    """Load a built shared library from flamegpu2_amd/lib; fail loudly if absent.

    Libraries are loaded with local symbol scope: every model library exports
    the same C entry points and must resolve them against itself."""
    p = lib_path(name)
    if not os.path.exists(p):
        raise NativeLibraryMissing(
            f"{p} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` first")
    return ctypes.CDLL(p, mode=ctypes.RTLD_LOCAL)
