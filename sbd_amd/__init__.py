"""sbd_amd -- MI355X-native engine for selected-basis (alpha x beta string) diagonalization.

The compute path is libsbdx.so (hand-written HIP kernels for gfx950 behind the C ABI in
include/sbdx.h); ``sbd_amd._lib`` binds it for Python callers, ``sbd_amd.synth`` builds the
seeded synthetic workloads.  Importing the package does not load the GPU library; creating an
``Engine`` does, and fails loudly if it has not been built.
"""
__all__ = ["Engine"]


def __getattr__(name):
    if name == "Engine":
        from ._lib import Engine
        return Engine
    raise AttributeError(name)
