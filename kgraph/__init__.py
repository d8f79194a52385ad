"""kgraph -- the module name the reference's Python binding registers
(python-api.cpp:412 PYBIND11_MODULE(kgraph, ...), setup.py:24), so existing
`import kgraph` code runs unchanged on the GPU implementation.

Module attributes as the reference's: `KGraph` (constructor, build / search /
load / save / save_text with the same keyword arguments and defaults),
`git_commit` and `arch` ({"name", "alignment"}).  Everything is the
kgraph_amd package (libkgraph_amd.so through its C ABI); this module only
gives it the reference's name.
"""
from kgraph_amd import KGraph, arch, git_commit, version  # noqa: F401

__all__ = ["KGraph", "git_commit", "arch", "version"]
