"""Field allocators (@Field, @XField, ..., @IField) without padding.

Sizes restate `_field` of src/ParallelKernel/FieldAllocators.jl:569-603
(padding=false branch): each staggered field is the grid size plus a fixed
per-dimension offset. Fields are pitch-aligned device arrays.
"""
from . import core

# size template -> offset added to the grid size (padding=false)
SIZE_OFFSETS = {
    None: (0, 0, 0),
    "X": (-1, -2, -2), "Y": (-2, -1, -2), "Z": (-2, -2, -1),
    "BX": (+1, 0, 0), "BY": (0, +1, 0), "BZ": (0, 0, +1),
    "XX": (0, -2, -2), "YY": (-2, 0, -2), "ZZ": (-2, -2, 0),
    "XY": (-1, -1, -2), "XZ": (-1, -2, -1), "YZ": (-2, -1, -1),
    "I": (-2, -2, -2),
    "XXY": (0, -1, -2), "XYY": (-1, 0, -2), "XYZ": (-1, -1, -1),
    "XXYZ": (0, -1, -1), "XYYZ": (-1, 0, -1), "XYZZ": (-1, -1, 0),
    "XXYY": (0, 0, -2), "XXZZ": (0, -2, 0), "YYZZ": (-2, 0, 0),
    "XXYYZ": (0, 0, -1), "XXYZZ": (0, -1, 0), "XYYZZ": (-1, 0, 0),
}


def field_size(gridsize, template=None):
    g = tuple(gridsize)
    off = SIZE_OFFSETS[template]
    return tuple(g[i] + off[i] for i in range(len(g)))


def _make(template):
    def alloc(gridsize, allocator=None, eltype=None):
        size = field_size(gridsize, template)
        if any(s < 0 for s in size):
            raise core.ArgumentError(f"grid size {tuple(gridsize)} too small for a {template or ''}Field")
        alloc_fn = allocator or core.zeros
        return alloc_fn(*size, eltype=eltype or core._numbertype_of(core._caller(1)))
    alloc.__name__ = f"{template or ''}Field"
    alloc.__doc__ = f"@{template or ''}Field(gridsize, allocator): array of size gridsize + {SIZE_OFFSETS[template]}."
    return alloc


Field = _make(None)
XField, YField, ZField = _make("X"), _make("Y"), _make("Z")
BXField, BYField, BZField = _make("BX"), _make("BY"), _make("BZ")
XXField, YYField, ZZField = _make("XX"), _make("YY"), _make("ZZ")
XYField, XZField, YZField = _make("XY"), _make("XZ"), _make("YZ")
IField = _make("I")
XXYField, XYYField, XYZField = _make("XXY"), _make("XYY"), _make("XYZ")
XXYZField, XYYZField, XYZZField = _make("XXYZ"), _make("XYYZ"), _make("XYZZ")
XXYYField, XXZZField, YYZZField = _make("XXYY"), _make("XXZZ"), _make("YYZZ")
XXYYZField, XXYZZField, XYYZZField = _make("XXYYZ"), _make("XXYZZ"), _make("XYYZZ")


def VectorField(gridsize, allocator=None, eltype=None):
    """(x=XField, y=YField, z=ZField) (FieldAllocators.jl:129-158)."""
    g = tuple(gridsize)
    if len(g) == 1:
        return {"x": XField(g, allocator, eltype)}
    if len(g) == 2:
        return {"x": XField(g, allocator, eltype), "y": YField(g, allocator, eltype)}
    return {"x": XField(g, allocator, eltype), "y": YField(g, allocator, eltype), "z": ZField(g, allocator, eltype)}


def BVectorField(gridsize, allocator=None, eltype=None):
    g = tuple(gridsize)
    names = ("x", "y", "z")[: len(g)]
    fns = (BXField, BYField, BZField)
    return {n: fns[i](g, allocator, eltype) for i, n in enumerate(names)}


def TensorField(gridsize, allocator=None, eltype=None):
    g = tuple(gridsize)
    if len(g) == 2:
        return {"xx": XXField(g, allocator, eltype), "yy": YYField(g, allocator, eltype), "xy": XYField(g, allocator, eltype)}
    return {"xx": XXField(g, allocator, eltype), "yy": YYField(g, allocator, eltype), "zz": ZZField(g, allocator, eltype),
            "xy": XYField(g, allocator, eltype), "xz": XZField(g, allocator, eltype), "yz": YZField(g, allocator, eltype)}
