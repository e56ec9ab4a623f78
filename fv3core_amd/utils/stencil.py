"""Index algebra for cell-centred fields with halos.

Restates the integer behaviour of the reference's ``GridIndexing`` and
``StencilFactory`` (fv3core/utils/stencil.py:280-728) without gt4py: origins,
domains, storage shapes, vertical restriction and the tile-edge anchors that
``axis_offsets`` hands to stencils as externals (stencil.py:457-495).

Where gt4py expresses an anchor as ``gtscript.I[0] + n`` or ``I[-1] + n``, this
module returns an :class:`AxisOffset` (``anchor`` is ``"start"`` or ``"end"`` of
the stencil's own domain).  ``absolute`` turns an anchor into an index of the
storage, which is what the HIP kernels and the oracle use.  A rank that is not
on an edge gets the reference's sentinel (``+-32767``, numpy int16 max) so that
the region never matches.
"""
from __future__ import annotations

import dataclasses
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

X_DIM = "x"
X_INTERFACE_DIM = "x_interface"
Y_DIM = "y"
Y_INTERFACE_DIM = "y_interface"
Z_DIM = "z"
Z_INTERFACE_DIM = "z_interface"
X_DIMS = (X_DIM, X_INTERFACE_DIM)
Y_DIMS = (Y_DIM, Y_INTERFACE_DIM)
Z_DIMS = (Z_DIM, Z_INTERFACE_DIM)

INT16_MAX = 32767

Index3D = Tuple[int, int, int]


@dataclasses.dataclass(frozen=True)
class AxisOffset:
    """An index relative to the start (``I[0]``) or end (``I[-1]``) of a domain."""

    axis: str  # "I" or "J"
    anchor: str  # "start" or "end"
    offset: int

    def __add__(self, n: int) -> "AxisOffset":
        return AxisOffset(self.axis, self.anchor, self.offset + n)

    def __sub__(self, n: int) -> "AxisOffset":
        return AxisOffset(self.axis, self.anchor, self.offset - n)

    def absolute(self, origin: int, extent: int) -> int:
        """Storage index of this anchor for a domain ``[origin, origin+extent)``."""
        if self.anchor == "start":
            return origin + self.offset
        return origin + extent - 1 + self.offset

    def __repr__(self):
        base = f"{self.axis}[0]" if self.anchor == "start" else f"{self.axis}[-1]"
        if self.offset == 0:
            return base
        return f"{base} {'+' if self.offset > 0 else '-'} {abs(self.offset)}"


class _Axis:
    def __init__(self, name):
        self._name = name

    def __getitem__(self, idx):
        if idx == 0:
            return AxisOffset(self._name, "start", 0)
        if idx == -1:
            return AxisOffset(self._name, "end", 0)
        raise IndexError("only I[0] and I[-1] anchors exist")


I = _Axis("I")  # noqa: E741
J = _Axis("J")


def _extent(dim: str, nx: int, ny: int, nz: int) -> int:
    if dim == X_DIM:
        return nx
    if dim == X_INTERFACE_DIM:
        return nx + 1
    if dim == Y_DIM:
        return ny
    if dim == Y_INTERFACE_DIM:
        return ny + 1
    if dim == Z_DIM:
        return nz
    if dim == Z_INTERFACE_DIM:
        return nz + 1
    raise ValueError(f"unknown dimension {dim}")


class GridIndexing:
    """Indices for cell-centred variables with halos (stencil.py:280-642)."""

    def __init__(
        self,
        domain: Index3D,
        n_halo: int,
        south_edge: bool,
        north_edge: bool,
        west_edge: bool,
        east_edge: bool,
    ):
        self.origin = (n_halo, n_halo, 0)
        self.n_halo = n_halo
        self.domain = tuple(domain)
        self.south_edge = bool(south_edge)
        self.north_edge = bool(north_edge)
        self.west_edge = bool(west_edge)
        self.east_edge = bool(east_edge)

    # -- sizes ------------------------------------------------------------
    @property
    def max_shape(self) -> Index3D:
        return self.domain_full(add=(1, 1, 1 + self.origin[2]))

    @property
    def isc(self):
        return self.origin[0]

    @property
    def iec(self):
        return self.origin[0] + self.domain[0] - 1

    @property
    def jsc(self):
        return self.origin[1]

    @property
    def jec(self):
        return self.origin[1] + self.domain[1] - 1

    @property
    def isd(self):
        return self.origin[0] - self.n_halo

    @property
    def ied(self):
        return self.isd + self.domain[0] + 2 * self.n_halo - 1

    @property
    def jsd(self):
        return self.origin[1] - self.n_halo

    @property
    def jed(self):
        return self.jsd + self.domain[1] + 2 * self.n_halo - 1

    @property
    def nw_corner(self):
        return self.north_edge and self.west_edge

    @property
    def sw_corner(self):
        return self.south_edge and self.west_edge

    @property
    def ne_corner(self):
        return self.north_edge and self.east_edge

    @property
    def se_corner(self):
        return self.south_edge and self.east_edge

    def origin_full(self, add: Index3D = (0, 0, 0)) -> Index3D:
        return (self.isd + add[0], self.jsd + add[1], self.origin[2] + add[2])

    def origin_compute(self, add: Index3D = (0, 0, 0)) -> Index3D:
        return (self.isc + add[0], self.jsc + add[1], self.origin[2] + add[2])

    def domain_full(self, add: Index3D = (0, 0, 0)) -> Index3D:
        return (
            self.ied + 1 - self.isd + add[0],
            self.jed + 1 - self.jsd + add[1],
            self.domain[2] + add[2],
        )

    def domain_compute(self, add: Index3D = (0, 0, 0)) -> Index3D:
        return (
            self.iec + 1 - self.isc + add[0],
            self.jec + 1 - self.jsc + add[1],
            self.domain[2] + add[2],
        )

    # -- tile-edge anchors -----------------------------------------------
    def axis_offsets(self, origin: Sequence[int], domain: Sequence[int]) -> Dict:
        """Region anchors relative to a stencil's origin/domain (stencil.py:457-495)."""
        if self.west_edge:
            i_start = I[0] + (self.origin[0] - origin[0])
        else:
            i_start = I[0] - INT16_MAX
        if self.east_edge:
            i_end = I[-1] + ((self.origin[0] + self.domain[0]) - (origin[0] + domain[0]))
        else:
            i_end = I[-1] + INT16_MAX
        if self.south_edge:
            j_start = J[0] + (self.origin[1] - origin[1])
        else:
            j_start = J[0] - INT16_MAX
        if self.north_edge:
            j_end = J[-1] + ((self.origin[1] + self.domain[1]) - (origin[1] + domain[1]))
        else:
            j_end = J[-1] + INT16_MAX
        return {
            "i_start": i_start,
            "local_is": I[0] + (self.isc - origin[0]),
            "i_end": i_end,
            "local_ie": I[-1] + (self.iec - origin[0] - domain[0] + 1),
            "j_start": j_start,
            "local_js": J[0] + (self.jsc - origin[1]),
            "j_end": j_end,
            "local_je": J[-1] + (self.jec - origin[1] - domain[1] + 1),
        }

    def edge_anchors(self) -> Dict[str, Optional[int]]:
        """Absolute storage indices of the tile edges, ``None`` when not on that edge.

        For every stencil these equal ``axis_offsets(...)[x].absolute(origin, domain)``
        (``i_start`` is always ``isc`` and ``i_end`` always ``iec``), which is
        how the kernels receive them.
        """
        return {
            "i_start": self.isc if self.west_edge else None,
            "i_end": self.iec if self.east_edge else None,
            "j_start": self.jsc if self.south_edge else None,
            "j_end": self.jec if self.north_edge else None,
        }

    # -- dims based helpers ---------------------------------------------
    def _origin_from_dims(self, dims: Iterable[str]) -> List[int]:
        out = []
        for dim in dims:
            if dim in X_DIMS:
                out.append(self.origin[0])
            elif dim in Y_DIMS:
                out.append(self.origin[1])
            elif dim in Z_DIMS:
                out.append(self.origin[2])
        return out

    def _extent_from_dims(self, dims: Iterable[str]) -> List[int]:
        nx, ny, nz = self.domain
        return [_extent(d, nx, ny, nz) for d in dims]

    def get_origin_domain(
        self, dims: Sequence[str], halos: Sequence[int] = tuple()
    ) -> Tuple[Tuple[int, ...], Tuple[int, ...]]:
        origin = self._origin_from_dims(dims)
        domain = self._extent_from_dims(dims)
        for i, n in enumerate(halos):
            origin[i] -= n
            domain[i] += 2 * n
        return tuple(origin), tuple(domain)

    def get_shape(self, dims: Sequence[str], halos: Sequence[int] = tuple()):
        shape = self._extent_from_dims(dims)
        for i, d in enumerate(dims):
            if d in X_DIMS + Y_DIMS:
                shape[i] += self.n_halo
        for i, n in enumerate(halos):
            shape[i] += n
        return tuple(shape)

    def restrict_vertical(self, k_start=0, nk=None) -> "GridIndexing":
        if k_start < 0:
            raise ValueError("k_start must be positive")
        if k_start > self.domain[2]:
            raise ValueError(
                "k_start must be less than the number of vertical levels "
                f"(received {k_start} for {self.domain[2]} vertical levels"
            )
        if nk is None:
            nk = self.domain[2] - k_start
        elif nk < 0:
            raise ValueError("number of vertical levels should be positive")
        elif nk > (self.domain[2] - k_start):
            raise ValueError(
                "nk can be at most the size of the vertical domain minus k_start"
            )
        new = GridIndexing(
            self.domain[:2] + (nk,),
            self.n_halo,
            self.south_edge,
            self.north_edge,
            self.west_edge,
            self.east_edge,
        )
        new.origin = self.origin[:2] + (self.origin[2] + k_start,)
        return new


@dataclasses.dataclass
class StencilConfig:
    """Execution options; the backend is always the HIP library."""

    backend: str = "hip"
    rebuild: bool = False
    validate_args: bool = True
    device_sync: bool = False


class StencilFactory:
    """Holds the grid indexing every component is built from (stencil.py:645-728)."""

    def __init__(self, config: StencilConfig, grid_indexing: GridIndexing):
        self.config = config
        self.grid_indexing = grid_indexing

    def restrict_vertical(self, k_start=0, nk=None) -> "StencilFactory":
        return StencilFactory(
            config=self.config,
            grid_indexing=self.grid_indexing.restrict_vertical(k_start=k_start, nk=nk),
        )
