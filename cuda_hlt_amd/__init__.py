"""MI355X-native VELO reconstruction (raw SuperPixel banks to consolidated
tracks) behind the reference's sequence / algorithm-handler API.

native    ctypes binding of the C ABI (include/velo_hlt.h)
sequence  the plain Velo sequence driven from PyTorch
gen       seeded synthetic events (include/velo_gen.h)
build     in-tree build of the native libraries
"""
__all__ = ["native", "sequence", "gen", "build"]
