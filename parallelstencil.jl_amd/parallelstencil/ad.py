"""Automatic differentiation of kernels: `@parallel ∇=(A->Ā, B->B̄) f!(A, B, a)`.

Mirrors src/ParallelKernel/parallel.jl:265-301 (parallel_call_ad, the high-level `∇` syntax)
and the generic form it expands to, `AD.autodiff_deferred!(Enzyme.Reverse, f!,
DuplicatedNoNeed(A, Ā), ..., Const(a))` (src/ParallelKernel/AD.jl,
src/ParallelKernel/EnzymeExt/autodiff_gpu.jl). The derivative is not obtained from Enzyme but
from a source transformation of the lowered kernel in libpsamd.so (ps_kernel_derivative): the
result is a kernel of its own, taking the primal arguments followed by one shadow array per
duplicated argument, launched with the primal's ranges.

Python spelling: `f(A, B, a, **{"∇": [(A, Ā), (B, B̄)]})`, or `grad=` for `∇`; the keys of
`∇` and the values of `ad_annotations` are the call's argument objects or the kernel's
parameter names. `ad_mode` is `Reverse` (default) or `Forward`.

Semantics (the same as Enzyme's for race-free kernels):
- Reverse mode: on entry the shadow of each written duplicated array holds the adjoint of its
  new value; the kernel consumes it (sets it to zero where the array is written) and adds the
  adjoints of everything read into the shadows of the read duplicated arrays.
- Forward mode: the shadows of read arrays hold tangents; the shadows of written arrays
  receive the tangents of the new values.
- The primal update of an array is performed for Const and Duplicated arrays and skipped for
  DuplicatedNoNeed ones. Active scalars are accepted: as in the reference, where the
  derivative runs inside each GPU thread, their derivative is not returned.
"""
import json

from ._lib import ArgumentError, IncoherentArgumentError, KeywordArgumentError

Reverse = "Reverse"
Forward = "Forward"

ERRMSG_RETURN_ACTIVITY = ("AD.autodiff_deferred!: explicit insertion of the return type activity (third argument) is not supported as "
                          "not GPU compiler compatible; call without it instead.")
ERRMSG_INCOHERENT_AD = ("incoherent arguments `ad_mode`/`ad_annotations` in @parallel call: AD keywords are only valid if automatic "
                        "differentiation is triggered with the keyword argument `∇`.")


class Annotation:
    """An argument with its activity (Enzyme.Const / Active / Duplicated / DuplicatedNoNeed)."""
    kind = None

    def __init__(self, val, dval=None):
        self.val = val
        self.dval = dval

    def __repr__(self):
        return f"{self.kind}({self.val!r}" + ("" if self.dval is None else f", {self.dval!r}") + ")"


class Const(Annotation):
    kind = "Const"


class Active(Annotation):
    kind = "Active"


class Duplicated(Annotation):
    kind = "Duplicated"


class DuplicatedNoNeed(Annotation):
    kind = "DuplicatedNoNeed"


SUPPORTED_ANNOTATIONS = {"Const": Const, "Active": Active, "Duplicated": Duplicated, "DuplicatedNoNeed": DuplicatedNoNeed}
AD_DUPLICATE_DEFAULT = DuplicatedNoNeed
AD_ANNOTATION_DEFAULT = Const


def _mode_name(mode):
    m = str(mode)
    if m.startswith("Enzyme."):
        m = m[len("Enzyme."):]
    if m not in (Reverse, Forward):
        raise KeywordArgumentError(f"ad_mode must be Enzyme.Reverse or Enzyme.Forward (obtained: {mode})")
    return m


def autodiff_deferred_(mode, f, *args, ranges=None, nblocks=None, nthreads=None, stream=None, async_=False, shmem=None):
    """AD.autodiff_deferred!(mode, f!, args...) launched over the ranges of the call f!(args...):
    arguments that are not annotations are promoted to Const. Passing the return activity as
    third argument (the class Const itself) raises an ArgumentError, as in the reference."""
    from .core import CellArray, DeviceArray, Kernel
    if not isinstance(f, Kernel):
        raise ArgumentError("AD.autodiff_deferred!: the second argument must be a kernel")
    if args and isinstance(args[0], type) and issubclass(args[0], Annotation):
        raise ArgumentError(ERRMSG_RETURN_ACTIVITY)
    mode = _mode_name(mode)
    ann = [a if isinstance(a, Annotation) else Const(a) for a in args]
    k, flat = f._resolve(tuple(a.val for a in ann))
    comps = []  # one annotation per kernel parameter: tuple arguments are split into components
    for a in ann:
        if isinstance(a, (Duplicated, DuplicatedNoNeed)) and a.dval is not None:
            comps.extend(_expand(type(a), a.val, a.dval))
        else:
            comps.extend(type(a)(x, a.dval) for x in _components(a.val))
    if len(comps) != len(k.params):
        raise ArgumentError(f"MethodError: kernel {k.name} takes {len(k.params)} arguments ({len(comps)} given)")
    if any(isinstance(x, CellArray) for a in comps for x in (a.val, a.dval)):
        raise ArgumentError(f"∇: kernel {k.name} has CellArray arguments; derivative kernels take plain arrays")
    annots, shadows = {}, []
    for name, a in zip(k.params, comps):
        if isinstance(a, (Duplicated, DuplicatedNoNeed)):
            if a.dval is None:
                raise KeywordArgumentError(f"variable {name} is annotated {a.kind} but has no duplicate")
            if not isinstance(a.val, DeviceArray) or not isinstance(a.dval, DeviceArray):
                raise ArgumentError(f"{a.kind}({name}, ...): both the variable and its duplicate must be arrays")
            if a.val.shape != a.dval.shape or a.val.dtype_code != a.dval.dtype_code:
                raise ArgumentError(f"{a.kind}({name}, ...): the duplicate must have the shape and element type of the variable "
                                    f"({a.val.shape} {a.val.eltype} vs {a.dval.shape} {a.dval.eltype})")
            annots[name] = a.kind
            shadows.append(a.dval)
        elif isinstance(a, Active) and isinstance(a.val, DeviceArray):
            raise ArgumentError(f"Active({name}): arrays are passed as Duplicated or DuplicatedNoNeed")
    d = k._derivative(mode, annots)
    return d(*flat, *shadows, ranges=ranges, nblocks=nblocks, nthreads=nthreads, stream=stream, async_=async_, shmem=shmem)


def _as_list(v):
    return list(v) if isinstance(v, (tuple, list)) else [v]


def _is_pair(v):
    return isinstance(v, (tuple, list)) and len(v) == 2 and not any(isinstance(x, (tuple, list, dict)) for x in v)


def _grad_pairs(grad):
    """[(variable, duplicate)] from ∇: a mapping, one pair or a sequence of pairs."""
    if isinstance(grad, dict):
        return list(grad.items())
    if _is_pair(grad):
        return [tuple(grad)]
    if isinstance(grad, (tuple, list)) and all(_is_pair(p) for p in grad):
        return [tuple(p) for p in grad]
    raise KeywordArgumentError("∇ expects pairs variable -> duplicate: {A: Ā}, (A, Ā) or [(A, Ā), (B, B̄)]")


def call_with_gradient(kernel, args, grad, ad_mode=None, ad_annotations=None, **launch):
    """The high-level `@parallel ∇=... ad_mode=... ad_annotations=... f!(args...)`."""
    mode = _mode_name(Reverse if ad_mode is None else ad_mode)
    params = kernel.info.get("def_params", kernel.params)

    def position(var, what):
        if isinstance(var, str):
            if var in params:
                return params.index(var)
            raise KeywordArgumentError(f"variable {var} in {what} is not a positional argument of the kernel call {kernel.name}.")
        for i, a in enumerate(args):
            if a is var:
                return i
        raise KeywordArgumentError(f"variable {var!r} in {what} is not a positional argument of the kernel call {kernel.name}.")

    dup = {}
    for var, dval in _grad_pairs(grad):
        dup[position(var, "∇")] = dval
    byvar = {}
    for kw, vars_ in (ad_annotations or {}).items():
        kw = kw[len("Enzyme."):] if kw.startswith("Enzyme.") else kw
        if kw not in SUPPORTED_ANNOTATIONS:
            raise KeywordArgumentError(f"annotation {kw} is not (yet) supported with high-level syntax; use the generic syntax calling "
                                       "directly `autodiff_deferred!`.")
        for var in _as_list(vars_):
            i = position(var, "ad_annotations")
            if i in byvar:
                raise KeywordArgumentError(f"variable {params[i] if i < len(params) else var} has more than one annotation. Nested "
                                           "annotations are not (yet) supported with high-level syntax; use the generic syntax calling "
                                           "directly `autodiff_deferred!`.")
            byvar[i] = SUPPORTED_ANNOTATIONS[kw]
    annotated = []
    for i, a in enumerate(args):
        cls = byvar.get(i, AD_DUPLICATE_DEFAULT if i in dup else AD_ANNOTATION_DEFAULT)
        name = params[i] if i < len(params) else str(i)
        if i in dup:
            if cls in (Const, Active):
                raise KeywordArgumentError(f"variable {name} has a duplicate in ∇ but is annotated {cls.kind}")
            annotated.append(cls(a, dup[i]))
        else:
            if cls in (Duplicated, DuplicatedNoNeed):
                raise KeywordArgumentError(f"variable {name} is annotated {cls.kind} but has no duplicate in ∇")
            annotated.append(cls(a))
    return autodiff_deferred_(mode, kernel, *annotated, **launch)


def _components(a):
    from .core import _tuple_names, _tuple_values
    return [a] if _tuple_names(a) is None else _tuple_values(a)


def _expand(cls, a, da):
    """Annotated components of a (possibly tuple) argument and its duplicate."""
    ca, cd = _components(a), _components(da)
    if len(ca) != len(cd):
        raise ArgumentError("∇: a tuple argument and its duplicate must have the same components")
    return [cls(x, y) for x, y in zip(ca, cd)]


def spec_json(mode, annots, form="scatter"):
    return json.dumps({"mode": mode, "annotations": annots, "form": form})


def incoherent():
    raise IncoherentArgumentError(ERRMSG_INCOHERENT_AD)
