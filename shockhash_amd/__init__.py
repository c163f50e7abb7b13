"""MI355X (gfx950) ShockHash construction engine.

`shockhash_amd.engine` binds the C ABI of libshockhash_amd.so (declared in include/shockhash_amd.h);
`shockhash_amd.host` binds libshockhash_host.so, the C++ class templates ShockHash / SIMDShockHash /
ShockHash2 that call the engine. Both libraries are built in-tree by __graft_entry__.build().
"""
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
BUILD_DIR = os.path.join(PKG_DIR, "_build")
ENGINE_LIB = os.path.join(BUILD_DIR, "libshockhash_amd.so")
HOST_LIB = os.path.join(BUILD_DIR, "libshockhash_host.so")
