"""Field allocators (@allocate, @Field, @XField, ..., @IField, @VectorField, @BVectorField,
@TensorField) of src/ParallelKernel/FieldAllocators.jl.

`_field` (FieldAllocators.jl:569-637) in both modes of the caller module's `padding`:
  * padding=false: each staggered field is an array of the grid size plus a fixed offset per
    dimension (SIZE_OFFSETS);
  * padding=true: each field is a view of a parent array; the parent has the grid size plus one
    cell in the staggered dimensions (PARENT_EXTRA) and the view the box VIEW_BOX gives, so that
    fields of one grid share the parents' index space. Kernels defined with padding=true address
    the parents (`A.parent`) and mask with the views' boxes (`A.indices`).
The parents are pitch-aligned device arrays; a view keeps its parent alive.
"""
from . import core

# size template -> offset added to the grid size (padding=false, FieldAllocators.jl:584-603)
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

# padding=true: cells added to the grid size for the parent array (FieldAllocators.jl:572-580)
PARENT_EXTRA = {t: (0, 0, 0) for t in SIZE_OFFSETS}
for _ts, _e in ((("X", "BX", "XYY", "XYYZZ"), (1, 0, 0)), (("Y", "BY", "XXY", "XXYZZ"), (0, 1, 0)),
                (("Z", "BZ", "XXYYZ"), (0, 0, 1)), (("XY", "XYZZ"), (1, 1, 0)), (("XZ", "XYYZ"), (1, 0, 1)),
                (("YZ", "XXYZ"), (0, 1, 1)), (("XYZ",), (1, 1, 1))):
    for _t in _ts:
        PARENT_EXTRA[_t] = _e

# padding=true: the view of the parent, as (first index, offset of the last index from the parent's
# extent) per dimension (FieldAllocators.jl:622-634)
VIEW_BOX = {}
for _ts, _box in ((("X", "Y", "Z", "XY", "XZ", "YZ", "I", "XYZ"), ((2, -1), (2, -1), (2, -1))),
                  (("XX", "XXY", "XXYZ"), ((1, 0), (2, -1), (2, -1))),
                  (("YY", "XYY", "XYYZ"), ((2, -1), (1, 0), (2, -1))),
                  (("ZZ", "XYZZ"), ((2, -1), (2, -1), (1, 0))),
                  (("XXYY", "XXYYZ"), ((1, 0), (1, 0), (2, -1))),
                  (("XXZZ", "XXYZZ"), ((1, 0), (2, -1), (1, 0))),
                  (("YYZZ", "XYYZZ"), ((2, -1), (1, 0), (1, 0))),
                  ((None, "BX", "BY", "BZ"), ((1, 0), (1, 0), (1, 0)))):
    for _t in _ts:
        VIEW_BOX[_t] = _box


def field_size(gridsize, template=None):
    """Size of the field (padding=false: the array; padding=true: the view, the same size)."""
    g = tuple(gridsize)
    off = SIZE_OFFSETS[template]
    return tuple(g[i] + off[i] for i in range(len(g)))


def parent_size(gridsize, template=None):
    """padding=true: size of the parent array."""
    g = tuple(gridsize)
    e = PARENT_EXTRA[template]
    return tuple(g[i] + e[i] for i in range(len(g)))


def view_ranges(gridsize, template=None):
    """padding=true: the view's index ranges in its parent (A.indices), as (first, last) pairs."""
    p = parent_size(gridsize, template)
    box = VIEW_BOX[template]
    return tuple((box[i][0], p[i] + box[i][1]) for i in range(len(p)))


_ALLOCATORS = None


def _allocator(allocator):
    global _ALLOCATORS
    if _ALLOCATORS is None:
        _ALLOCATORS = {"zeros": core.zeros, "ones": core.ones, "rand": core.rand, "falses": core.falses, "trues": core.trues}
    if allocator is None:
        return core.zeros
    if isinstance(allocator, str):
        a = _ALLOCATORS.get(allocator.lstrip("@"))
        if a is None:
            raise core.ArgumentError("the allocator must be a field allocator (@zeros, @ones, @rand, @falses or @trues).")
        return a
    if allocator not in _ALLOCATORS.values():
        raise core.ArgumentError("the allocator must be a field allocator (@zeros, @ones, @rand, @falses or @trues).")
    return allocator


def _field(mod, gridsize, template, allocator, eltype):
    cfg = core._config(mod)
    eltype = eltype or core._numbertype_of(mod)
    alloc = _allocator(allocator)
    size = field_size(gridsize, template)
    if any(s < 0 for s in size):
        raise core.ArgumentError(f"grid size {tuple(gridsize)} too small for a {template or ''}Field")
    if not cfg.padding:
        return alloc(*size, eltype=eltype)
    parent = alloc(*parent_size(gridsize, template), eltype=eltype)
    rng = view_ranges(gridsize, template)
    return parent.view([r[0] for r in rng], [r[1] for r in rng])


def _make(template):
    def alloc(gridsize, allocator=None, eltype=None):
        return _field(core._caller(1), gridsize, template, allocator, eltype)
    alloc.__name__ = f"{template or ''}Field"
    alloc.__doc__ = (f"@{template or ''}Field(gridsize, allocator; eltype): field of size gridsize + {SIZE_OFFSETS[template]} "
                     f"(padding=true: a view of a parent of size gridsize + {PARENT_EXTRA[template]}).")
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


def _vector(mod, gridsize, templates, allocator, eltype):
    g = tuple(gridsize)
    names = ("x", "y", "z")[: len(g)]
    return {n: _field(mod, g, templates[i], allocator, eltype) for i, n in enumerate(names)}


def VectorField(gridsize, allocator=None, eltype=None):
    """(x=XField, y=YField, z=ZField) (FieldAllocators.jl:129-158)."""
    return _vector(core._caller(1), gridsize, ("X", "Y", "Z"), allocator, eltype)


def BVectorField(gridsize, allocator=None, eltype=None):
    """(x=BXField, y=BYField, z=BZField)."""
    return _vector(core._caller(1), gridsize, ("BX", "BY", "BZ"), allocator, eltype)


def TensorField(gridsize, allocator=None, eltype=None):
    """(xx, yy, xy) in 2-D, (xx, yy, zz, xy, xz, yz) in 3-D."""
    mod, g = core._caller(1), tuple(gridsize)
    names = ("xx", "yy", "xy") if len(g) == 2 else ("xx", "yy", "zz", "xy", "xz", "yz")
    return {n: _field(mod, g, n.upper(), allocator, eltype) for n in names}


FIELDTYPES = {"Field": None, "XField": "X", "YField": "Y", "ZField": "Z", "BXField": "BX", "BYField": "BY",
              "BZField": "BZ", "XXField": "XX", "YYField": "YY", "ZZField": "ZZ", "XYField": "XY", "XZField": "XZ",
              "YZField": "YZ", "VectorField": "vector", "BVectorField": "bvector", "TensorField": "tensor"}


def allocate(gridsize=None, fields=None, allocator=None, eltype=None, **kw):
    """@allocate(gridsize=..., fields=(FieldType => names, ...), allocator=@zeros, eltype=numbertype)
    (FieldAllocators.jl:66-72,535-567): allocates every named field on the same grid and returns
    them by name (Julia binds them as variables of the caller). `fields` is a dict or a sequence of
    (field type, name or tuple of names) pairs; field types are given by name ("XField") or as the
    allocator functions themselves (XField)."""
    if kw:
        raise core.ArgumentError(f"invalid keyword argument(s) in @allocate: {', '.join(sorted(kw))}")
    if gridsize is None or fields is None:
        raise core.ArgumentError("the gridsize and the fields keyword argument are mandatory.")
    _allocator(allocator)
    mod = core._caller(1)
    pairs = list(fields.items()) if isinstance(fields, dict) else list(fields)
    if pairs and isinstance(pairs[0], str):  # a single pair
        pairs = [tuple(pairs)]
    by_fn = {globals()[n]: n for n in FIELDTYPES}
    out = {}
    for T, names in pairs:
        tname = by_fn.get(T, T)
        if tname not in FIELDTYPES:
            raise core.ArgumentError(f"invalid field type in @allocate: {T} (valid: {', '.join(FIELDTYPES)}).")
        t = FIELDTYPES[tname]
        for name in ((names,) if isinstance(names, str) else tuple(names)):
            if t == "vector":
                out[name] = _vector(mod, gridsize, ("X", "Y", "Z"), allocator, eltype)
            elif t == "bvector":
                out[name] = _vector(mod, gridsize, ("BX", "BY", "BZ"), allocator, eltype)
            elif t == "tensor":
                g = tuple(gridsize)
                tn = ("xx", "yy", "xy") if len(g) == 2 else ("xx", "yy", "zz", "xy", "xz", "yz")
                out[name] = {n: _field(mod, g, n.upper(), allocator, eltype) for n in tn}
            else:
                out[name] = _field(mod, gridsize, t, allocator, eltype)
    return out
