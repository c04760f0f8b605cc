"""torch.autograd.gradcheck through best_hydro_amd.autograd (the adjoint as a torch.autograd.Function)
on the 80-triangle mesh.  It is the one GPU test that imports torch, and it sits after every test
that captures a HIP graph: once torch's own HIP libraries are loaded beside the library's, graph
capture is not run in the same process (see conftest.no_torch_loaded)."""
import numpy as np
import pytest

from test_gradient import tiny_mesh


@pytest.mark.gpu
def test_autograd_gradcheck(bh):
    import torch
    from best_hydro_amd.autograd import MeshTransport, transport
    mt = MeshTransport(tiny_mesh(1), precision=64, wm=14000.0)
    V = torch.tensor(mt.V, dtype=torch.float64, requires_grad=True)
    q0 = torch.tensor(mt.values(mt.V), dtype=torch.float64)
    f = lambda X: transport(X, mt) / q0
    assert torch.autograd.gradcheck(f, (V,), eps=1e-6, atol=1e-7, rtol=1e-5)
    # backward of a scalar combination equals the combination of the rows
    y = (transport(V, mt) * torch.tensor([1.0, 2.0, 3.0, 4.0], dtype=torch.float64) / q0).sum()
    y.backward()
    _, G = mt.values_and_gradient(mt.V)
    want = np.einsum("o,ovc->vc", np.array([1.0, 2.0, 3.0, 4.0]) / q0.numpy(), G)
    assert np.allclose(V.grad.numpy(), want, rtol=1e-12, atol=0)
    mt.solver.close()
