"""Gauge fields of gpt_amd: four SU(3) link fields in HBM.

Links are FP64 complex in natural site order [V][3][3] (site index
((t*Lz + z)*Ly + y)*Lx + x), the layout the operator's set-up reads
(include/gpt_amd_wc.h); boundary phases are not stored in them
(lib/gpt/core/covariant.py:28-37 applies them inside the operator).
"""
import numpy as np
import torch

from .. import core


class link_field:
    """One colour-matrix field U_mu (gpt's ot_matrix_su_n_fundamental_group(3))."""

    def __init__(self, grid, data=None):
        self.grid = grid
        if data is None:
            data = torch.zeros((grid.gsites, 3, 3), dtype=torch.complex128, device=core.device())
        self.data = data


def from_numpy(grid, U):
    """Four natural-layout arrays [t, z, y, x, 3, 3] -> list of link_field."""
    shape = tuple(reversed(grid.fdimensions)) + (3, 3)
    out = []
    for u in U:
        u = np.asarray(u)
        if u.shape != shape:
            raise ValueError(f"expected links of shape {shape}, got {u.shape}")
        t = torch.from_numpy(np.ascontiguousarray(u.reshape(grid.gsites, 3, 3), dtype=np.complex128))
        out.append(link_field(grid, t.to(core.device())))
    return out


def to_numpy(U):
    shape = tuple(reversed(U[0].grid.fdimensions)) + (3, 3)
    return [u.data.cpu().numpy().reshape(shape) for u in U]


def unit(grid):
    eye = torch.eye(3, dtype=torch.complex128, device=core.device())
    return [link_field(grid, eye.expand(grid.gsites, 3, 3).contiguous()) for _ in range(4)]


def stacked(U):
    """[4][V][3][3] complex128 contiguous device tensor for the operator set-up."""
    return torch.stack([u.data for u in U]).contiguous()


def plaquette(U):
    """g.qcd.gauge.plaquette (lib/gpt/qcd/gauge/loops.py): the average of
    Re tr of the plaquettes over the 6 planes, normalised to 1 for unit links,
    computed on the GPU (wc_gauge_observables)"""
    from .. import io

    return io.gauge_observables(U)[0]


def link_trace(U):
    """the average Re tr U_mu(x) / 3 over sites and directions, on the GPU"""
    from .. import io

    return io.gauge_observables(U)[1]
