"""FiniteDifferences2D: the macros available in 2-D @parallel kernels.

The macros are expanded by the kernel lowering in libpsamd.so (index algebra of
src/FiniteDifferences.jl); this module lists them for discovery and documentation.
"""
NDIMS = 2
MACROS = ("d_xa", "d_ya", "d_xi", "d_yi", "d2_xa", "d2_ya", "d2_xi", "d2_yi", "all", "inn", "inn_x", "inn_y",
          "av", "av_xa", "av_ya", "av_xi", "av_yi", "harm", "harm_xa", "harm_ya", "harm_xi", "harm_yi", "maxloc", "minloc")
ASSIGNABLE = ("all", "inn", "inn_x", "inn_y")
