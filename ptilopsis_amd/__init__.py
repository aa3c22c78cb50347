"""MI355X-native drop-in for Ptilopsis's data-parallel RISC-V code generator.

The product is the C-ABI library ``libptl_hip.so`` (declared in
``include/ptl.h``) built from ``ptilopsis_amd/csrc``; this package holds the
host-side mirror of the reference's ``rv_generator`` interface
(:mod:`ptilopsis_amd.generator`), the ctypes bindings, the front end used to
turn Pareas-language source into the generator's input arrays, and the named
synthetic workloads.
"""
from .native import host_lib, hip_lib, Tree, Arena, parse_source, parse_arena, generate_source  # noqa: F401
