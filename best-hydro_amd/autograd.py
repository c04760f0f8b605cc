"""D_t, D_r and CHI V (and [eta]) of a triangulation as differentiable functions of its shared
vertex coordinates: a ``torch.autograd.Function`` whose forward is one GPU solve and whose
backward is the adjoint gradient of ``bhg_gradient`` (include/best_hip.h).

The surfaces (COMPON's ``icomp``, best/compon.f90:19-115) are taken from the unperturbed mesh and
held fixed, so a vertex moved off its neighbours' copies does not change the deflation.

torch is imported only when ``transport`` is first called: ``MeshTransport`` (values and adjoint
gradients as numpy arrays) does not load it, so a process that only solves never carries torch's
own HIP libraries beside the library's.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import Solver, lib, meshes, _col


def fixed_icomp(tri) -> np.ndarray:
    """icomp as bhg_set_mesh would derive it (best/best.f:324-337): COMPON's surfaces, or one
    surface for all if any edge is open."""
    n = tri.nar
    icomp = np.zeros(n, dtype=np.int32)
    nc, no = C.c_int(), C.c_int()
    if lib().bhg_compon(n, _col(tri.v1), _col(tri.v2), _col(tri.v3), icomp, C.byref(nc), C.byref(no)) != 0:
        raise RuntimeError("bhg_compon failed")
    if no.value > 0:
        icomp[:] = 1
    return icomp


class MeshTransport:
    """A triangulation's connectivity, surfaces and header with a solver bound to it."""

    def __init__(self, tri, solver: Solver | None = None, iarea=1, cg=1.0, precision=64, wm=None):
        self.tri = tri
        self.V, self.idx = meshes.vertex_index(tri)
        self.icomp = fixed_icomp(tri)
        self.solver = solver or Solver()
        self.iarea, self.cg, self.precision, self.wm = iarea, cg, precision, wm

    def mesh(self, V):
        return meshes.with_vertices(self.tri, V, self.idx)

    def values(self, V):
        """(D_t, D_r, CHI V[, eta]) at vertices V by one solve."""
        return self.values_of(self.mesh(V))

    def values_of(self, tri):
        """The same outputs of a triangulation with this one's connectivity (surfaces held)."""
        self.solver.set_mesh(tri, icomp=self.icomp)
        r = self.solver.solve_mesh(self.iarea, self.cg, 1, 0, self.precision, self.wm)
        return self._pick(r)

    def values_and_gradient(self, V):
        """Values and d(values)/dV (outputs x vertices x 3) by one solve and its adjoint."""
        self.solver.set_mesh(self.mesh(V), icomp=self.icomp)
        r, g = self.solver.gradient(self.iarea, self.cg, self.precision, self.wm)
        return self._pick(r), meshes.vertex_gradient(g, self.idx, len(self.V))

    def _pick(self, r):
        keys = ["dtrans", "dr", "chiprime"] + (["eta"] if self.wm is not None else [])
        return np.array([r[k] for k in keys])


_FUNCTION = None


def transport_function():
    """The torch.autograd.Function: forward(V, mt) -> (D_t, D_r, CHI V[, eta]) as a float64 tensor,
    backward by the adjoint (built on first use, with torch imported then)."""
    global _FUNCTION
    if _FUNCTION is None:
        import torch

        class TransportFunction(torch.autograd.Function):
            @staticmethod
            def forward(ctx, V, mt: MeshTransport):
                vals, grad = mt.values_and_gradient(V.detach().cpu().numpy().astype(np.float64))
                ctx.save_for_backward(torch.from_numpy(grad))
                return torch.from_numpy(vals).to(V.device)

            @staticmethod
            def backward(ctx, gout):
                (grad,) = ctx.saved_tensors
                gv = torch.einsum("o,ovc->vc", gout.detach().cpu().to(torch.float64), grad)
                return gv.to(gout.device), None

        _FUNCTION = TransportFunction
    return _FUNCTION


def transport(V, mt: MeshTransport):
    """(D_t, D_r, CHI V[, eta]) of the vertices V (a torch tensor), differentiable in V."""
    return transport_function().apply(V, mt)
