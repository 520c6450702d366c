"""Host-side mirror of Octo-Tiger's interaction-kernel interfaces.

Same names, argument meanings and output conventions as the reference:

  multipole_interaction_interface.compute_multipole_interactions(
      monopoles, M_ptr, com_ptr, neighbors, type, dx, is_direction_empty, xbase,
      use_root_stencil)
    (octotiger/multipole_interactions/legacy/multipole_interaction_interface.hpp:33-77)
  monopole_interaction_interface.compute_interactions(
      monopoles, com_ptr, neighbors, type, dx, is_direction_empty, grid_ptr,
      contains_multipole_neighbor)
    (octotiger/monopole_interactions/legacy/monopole_interaction_interface.hpp:29-61)

Results go into the grid's L (accumulated) and L_c (overwritten), as the
reference writes grid::get_L() / grid::get_L_c().  Everything runs through
liboctofmm on the GPU.
"""
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _abi
from .native import Context

RHO = _abi.RHO
DRHODT = _abi.DRHODT


@dataclass
class NeighborData:
    """neighbor_gravity_type::data -- gravity_boundary_type (interaction_types.hpp:27-45)."""
    M: Optional[np.ndarray] = None      # [n, 20] multipoles
    m: Optional[np.ndarray] = None      # [n] monopoles
    x: Optional[np.ndarray] = None      # [n, 3] centres of mass
    cells: Optional[np.ndarray] = None  # partial halo: neighbour-local flat indices (ilist_n_bnd order)


@dataclass
class NeighborGravity:
    """neighbor_gravity_type (interaction_types.hpp:47-64)."""
    is_monopole: bool = False
    data: NeighborData = None
    direction: int = 0

    def __post_init__(self):
        if self.data is None:
            self.data = NeighborData()


class Grid:
    """The parts of grid that the kernels write: L and L_c (512 cells)."""

    def __init__(self):
        self.L = np.zeros((512, 20))
        self.L_c = np.zeros((512, 3))

    def get_L(self):
        return self.L

    def get_L_c(self):
        return self.L_c


def _f64(a, shape):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.shape != shape and a.size != int(np.prod(shape)):
        raise ValueError(f"expected {shape}, got {a.shape}")
    return a.reshape(shape)


def _neighbor_struct(nb, keep):
    s = _abi.Neighbor()
    d = nb.data
    if not nb.is_monopole:
        if d.M is None or d.x is None:
            s.kind = _abi.NEIGHBOR_EMPTY
            return s
        M = np.ascontiguousarray(d.M, dtype=np.float64).reshape(-1, 20)
        x = np.ascontiguousarray(d.x, dtype=np.float64).reshape(-1, 3)
        n = M.shape[0]
        s.kind = _abi.NEIGHBOR_MULTIPOLE
        s.M, s.com = _abi.dptr(M), _abi.dptr(x)
        keep += [M, x]
    else:
        if d.m is None:
            s.kind = _abi.NEIGHBOR_EMPTY
            return s
        m = np.ascontiguousarray(d.m, dtype=np.float64).reshape(-1)
        n = m.shape[0]
        s.kind = _abi.NEIGHBOR_MONOPOLE
        s.mon = _abi.dptr(m)
        keep.append(m)
    s.count = n
    if n != 512 or d.cells is not None:
        if d.cells is None:
            raise ValueError("a partial halo needs its cell list")
        cells = np.ascontiguousarray(d.cells, dtype=np.int32)
        if cells.size != n:
            raise ValueError("cell list and data sizes differ")
        s.cells = _abi.iptr(cells)
        keep.append(cells)
    return s


def _subgrid(monopoles, M_ptr, com_ptr, neighbors, type_, dx, xbase, root):
    keep = []
    sg = _abi.Subgrid()
    if M_ptr is not None and len(M_ptr):
        M = _f64(M_ptr, (512, 20))
        keep.append(M)
        sg.M = _abi.dptr(M)
    com = _f64(com_ptr[0], (512, 3))
    keep.append(com)
    sg.com = _abi.dptr(com)
    if monopoles is not None and len(monopoles):
        mon = _f64(monopoles, (512,))
        keep.append(mon)
        sg.mon = _abi.dptr(mon)
    if len(neighbors) != 26:
        raise ValueError("26 neighbours expected (geo::direction flat order)")
    for f, nb in enumerate(neighbors):
        sg.neighbors[f] = _neighbor_struct(nb, keep)
    sg.dx = float(dx)
    for k in range(3):
        sg.xbase[k] = float(xbase[k])
    sg.type = int(type_)
    sg.use_root_stencil = 1 if root else 0
    return sg, keep


class multipole_interaction_interface:
    """Non-leaf sub-grids: M2L over the opening-criterion stencil, or over all
    pairs of the root sub-grid."""

    def __init__(self, context: Context = None, theta=0.5, device=0):
        self.ctx = context or Context(theta=theta, device=device)
        self.grid_ptr = None

    def set_grid_ptr(self, grid):
        self.grid_ptr = grid

    def compute_multipole_interactions(self, monopoles, M_ptr, com_ptr, neighbors, type, dx,
                                       is_direction_empty, xbase, use_root_stencil):
        sg, keep = _subgrid(monopoles, M_ptr, com_ptr, neighbors, type, dx, xbase, use_root_stencil)
        g = self.grid_ptr
        self.ctx.multipole_interactions(sg, g.get_L(), g.get_L_c())


class monopole_interaction_interface:
    """Leaf sub-grids: p2p against monopoles, p2m against refined neighbours."""

    def __init__(self, context: Context = None, theta=0.5, device=0):
        self.ctx = context or Context(theta=theta, device=device)

    def compute_interactions(self, monopoles, com_ptr, neighbors, type, dx, is_direction_empty, grid_ptr,
                             contains_multipole_neighbor, xbase=(0.0, 0.0, 0.0)):
        sg, keep = _subgrid(monopoles, None, com_ptr, neighbors, type, dx, xbase, False)
        self.ctx.monopole_interactions(sg, contains_multipole_neighbor, grid_ptr.get_L(), grid_ptr.get_L_c())


def neighbors_from_level(level, slot):
    """NeighborGravity list (flat direction order) of one slot of a
    workloads.Level."""
    out = []
    for f in range(26):
        raw = _abi.flat_to_raw(f)
        n = int(level.neighbors[slot, raw])
        if n < 0:
            out.append(NeighborGravity(is_monopole=False, direction=f))
        elif level.kind[n] == _abi.SLOT_LEAF:
            out.append(NeighborGravity(is_monopole=True, data=NeighborData(m=level.mon[n]), direction=f))
        else:
            out.append(NeighborGravity(is_monopole=False,
                                       data=NeighborData(M=level.M[n], x=level.com[n]), direction=f))
    return out
