"""ImplicitGlobalGrid on one GPU: the grid functions the reference's multi-XPU examples call
(examples/diffusion3D_multigpucpu_hidecomm*.jl, miniapps/acoustic_waves_multixpu) for a single
process.

ImplicitGlobalGrid is a dependency of those examples, not part of the reference
(ImplicitGlobalGrid.jl 0.15, not vendored); what is restated here is its published single-process
behaviour:
- the process grid is 1 x 1 x 1 (me = 0, coords = (0, 0, 0));
- the global size of a dimension is nx - overlap when it is periodic, nx otherwise;
- x_g(ix, dx, A) is the global coordinate of index ix of A: staggered arrays (size nx + k) are
  shifted by -k dx / 2, and periodic dimensions map the halo cells onto their periodic images;
- update_halo!(A...) exchanges the halo of A with the neighbours of each dimension. With one
  process the only neighbour is the process itself, through a periodic boundary: the halo cells
  1:hw take the cells s-ol+1 : s-ol+hw and the cells s-hw+1:s take ol-hw+1 : ol (s the size of A,
  ol its overlap, nx overlap + s - nx, hw the halo width), dimension after dimension, so that
  edges and corners are filled too. Non-periodic dimensions have no neighbour: nothing happens.

The halo copies are device-to-device copies between views of the array (the packing and
unpacking that a multi-GPU exchange would do into its send and receive buffers).
"""
import time

import numpy as np

from . import core
from ._lib import ArgumentError

_grid = None


class _Grid:
    def __init__(self, nxyz, periods, overlaps, halowidths):
        self.nxyz = nxyz
        self.periods = periods
        self.overlaps = overlaps
        self.halowidths = halowidths
        self.dims = (1, 1, 1)
        self.coords = (0, 0, 0)
        self.me = 0
        self.nprocs = 1


def _g():
    if _grid is None:
        raise ArgumentError("ImplicitGlobalGrid: no function of the global grid can be called before init_global_grid.")
    return _grid


def init_global_grid(nx, ny=1, nz=1, *, dimx=0, dimy=0, dimz=0, periodx=0, periody=0, periodz=0, overlaps=(2, 2, 2), halowidths=None,
                     disp=1, reorder=1, comm=None, init_MPI=True, device_type="AMDGPU", select_device=True, quiet=False):
    """(me, dims, nprocs, coords, comm) of the one-process grid."""
    global _grid
    if _grid is not None:
        raise ArgumentError("ImplicitGlobalGrid: the global grid has already been initialized.")
    for name, d in (("dimx", dimx), ("dimy", dimy), ("dimz", dimz)):
        if d not in (0, 1):
            raise ArgumentError(f"ImplicitGlobalGrid: {name}={d}: this backend runs one process on one GPU (dims 1 x 1 x 1).")
    nxyz = (int(nx), int(ny), int(nz))
    overlaps = tuple(int(o) for o in overlaps)
    halowidths = tuple(max(1, o // 2) for o in overlaps) if halowidths is None else tuple(int(h) for h in halowidths)
    periods = (int(bool(periodx)), int(bool(periody)), int(bool(periodz)))
    for d in range(3):
        if nxyz[d] > 1 and periods[d] and nxyz[d] < 2 * overlaps[d] - 1:
            raise ArgumentError(f"ImplicitGlobalGrid: the local size in dimension {d + 1} ({nxyz[d]}) is too small for overlap {overlaps[d]}.")
        if halowidths[d] < 1 or (overlaps[d] > 0 and halowidths[d] > overlaps[d] // 2 and nxyz[d] > 1):
            raise ArgumentError(f"ImplicitGlobalGrid: halowidth {halowidths[d]} in dimension {d + 1} must be >= 1 and <= overlap/2.")
    _grid = _Grid(nxyz, periods, overlaps, halowidths)
    if select_device:
        core.ensure_device()
    return 0, (1, 1, 1), 1, (0, 0, 0), None


def finalize_global_grid(finalize_MPI=True):
    global _grid
    _g()
    _grid = None
    _halo_views.clear()


def select_device():
    """The GPU of this process (device 0)."""
    core.ensure_device()
    return 0


def _n_g(d):
    g = _g()
    return g.dims[d] * (g.nxyz[d] - g.overlaps[d]) + g.overlaps[d] * (g.periods[d] == 0)


def nx_g():
    return _n_g(0)


def ny_g():
    return _n_g(1)


def nz_g():
    return _n_g(2)


def _coord_g(d, i, dxy, A):
    g = _g()
    size = A.shape[d] if d < len(A.shape) else 1
    x0 = 0.5 * (g.nxyz[d] - size) * dxy
    x = (g.coords[d] * (g.nxyz[d] - g.overlaps[d]) + i - 1) * dxy + x0
    if g.periods[d]:
        n = _n_g(d)
        x = x - dxy
        if x > (n - 1) * dxy:
            x = x - n * dxy
        if x < 0:
            x = x + n * dxy
    return x


def x_g(ix, dx, A):
    return _coord_g(0, ix, dx, A)


def y_g(iy, dy, A):
    return _coord_g(1, iy, dy, A)


def z_g(iz, dz, A):
    return _coord_g(2, iz, dz, A)


def _halo_boxes(shape, d, ol, hw):
    """[(destination box, source box)] of the periodic self-exchange of dimension d (1-based,
    inclusive boxes over the whole extent of the other dimensions)."""
    s = shape[d]

    def box(a, b):
        lo = [1, 1, 1]
        hi = list(shape) + [1] * (3 - len(shape))
        lo[d], hi[d] = a, b
        return tuple(lo), tuple(hi)

    return [(box(1, hw), box(s - ol + 1, s - ol + hw)), (box(s - hw + 1, s), box(ol - hw + 1, ol))]


_halo_views = {}  # id(A) -> (A, [(destination view, source view)]): prepared once per array


def _halo_pairs(A, g):
    ent = _halo_views.get(id(A))
    if ent is not None and ent[0] is A:
        return ent[1]
    shape = tuple(A.shape) + (1,) * (3 - len(A.shape))
    pairs = []
    for d in range(3):
        if not g.periods[d] or g.nxyz[d] <= 1 or shape[d] <= 1:
            continue
        ol = g.overlaps[d] + shape[d] - g.nxyz[d]
        hw = g.halowidths[d]
        if ol < 2 * hw:
            raise ArgumentError(f"update_halo!: the overlap of the array in dimension {d + 1} ({ol}) is smaller than 2 x halowidth ({hw}).")
        for dst, src in _halo_boxes(shape, d, ol, hw):
            pairs.append((A.view(*dst), A.view(*src)))
    if len(_halo_views) >= 64:
        _halo_views.pop(next(iter(_halo_views)))
    _halo_views[id(A)] = (A, pairs)
    return pairs


def update_halo_(*arrays, stream=None, async_=False):
    """update_halo!(A...): the halo exchange of each array (periodic dimensions exchange with
    the process itself). Ordered on `stream` (the current stream by default) and, as in
    ImplicitGlobalGrid, complete when it returns unless async_ is set."""
    g = _g()
    for A in arrays:
        if not isinstance(A, core.DeviceArray) or isinstance(A, core.CellArray):
            raise ArgumentError("update_halo!: the arguments must be fields (device arrays)")
        for dst, src in _halo_pairs(A, g):  # dimension after dimension: edges and corners follow
            dst.copy_(src, stream=stream, async_=True)
    if not async_:
        core.check(core.lib.ps_stream_sync(core._stream_handle(stream)))


def gather_(A, A_global, root=0):
    """gather!(A, A_global): with one process, A_global receives A (a host array)."""
    h = A.to_host()
    if A_global is None:
        return None
    if tuple(np.shape(A_global)) != tuple(h.shape):
        raise ArgumentError(f"gather!: A_global must have the size of A on one process (obtained {np.shape(A_global)} vs {h.shape})")
    A_global[...] = h
    return None


_t0 = None


def tic():
    global _t0
    core.synchronize()
    _t0 = time.perf_counter()


def toc():
    core.synchronize()
    return time.perf_counter() - _t0
