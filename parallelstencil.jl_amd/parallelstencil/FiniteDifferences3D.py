"""FiniteDifferences3D: the macros available in 3-D @parallel kernels.

The macros are expanded by the kernel lowering in libpsamd.so (index algebra of
src/FiniteDifferences.jl); this module lists them for discovery and documentation.
"""
NDIMS = 3
MACROS = ("d_xa", "d_ya", "d_za", "d_xi", "d_yi", "d_zi", "d2_xi", "d2_yi", "d2_zi", "all", "inn", "inn_x", "inn_y",
          "inn_z", "inn_xy", "inn_xz", "inn_yz", "av", "av_xa", "av_ya", "av_za", "av_xi", "av_yi", "av_zi", "av_xya",
          "av_xza", "av_yza", "av_xyi", "av_xzi", "av_yzi", "harm", "harm_xa", "harm_ya", "harm_za", "harm_xi",
          "harm_yi", "harm_zi", "harm_xya", "harm_xza", "harm_yza", "harm_xyi", "harm_xzi", "harm_yzi", "maxloc",
          "minloc")
ASSIGNABLE = ("all", "inn", "inn_x", "inn_y", "inn_z", "inn_xy", "inn_xz", "inn_yz")
