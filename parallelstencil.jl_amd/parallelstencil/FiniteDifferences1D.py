"""FiniteDifferences1D: the macros available in 1-D @parallel kernels.

The macros are expanded by the kernel lowering in libpsamd.so (index algebra of
src/FiniteDifferences.jl); this module lists them for discovery and documentation.
"""
NDIMS = 1
MACROS = ("d", "d2", "all", "inn", "av", "harm", "maxloc", "minloc")
ASSIGNABLE = ("all", "inn")
