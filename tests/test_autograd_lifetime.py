"""Lifetime of a projector under autograd graphs, on the CPU: a graph built by
autograd.project holds its projector, and close() on a projector that a graph
still uses only marks it closed; the context is released with the last graph.
The context itself is faked (no device): only the Python bookkeeping runs."""
import ctypes as C
import gc

import pytest

import ct_forward_projection_amd as P
from ct_forward_projection_amd import _lib as L


def fake_projector():
    proj = P.Projector.__new__(P.Projector)
    proj._ctx = C.c_void_p(0x1000)
    proj._graphs = 0
    proj._close_pending = False
    proj.destroyed = 0

    def destroy():
        proj.destroyed += 1
        proj._ctx = C.c_void_p()

    proj._destroy = destroy
    return proj


def test_close_without_graph_releases_at_once():
    proj = fake_projector()
    proj.close()
    assert proj.destroyed == 1
    proj.close()
    assert proj.destroyed == 1
    with pytest.raises(L.CTFPError) as e:
        proj.handle
    assert e.value.code == L.CTFP_ERR_STATE


def test_close_waits_for_graph_tokens():
    proj = fake_projector()
    t1, t2 = proj._graph_token(), proj._graph_token()
    proj.close()
    assert proj.destroyed == 0
    with pytest.raises(L.CTFPError):
        proj.handle  # closed for new calls
    del t1
    gc.collect()
    assert proj.destroyed == 0
    del t2
    gc.collect()
    assert proj.destroyed == 1


def test_graph_node_holds_the_projector():
    """The identity node that autograd.project puts in front of the
    extension holds the projector and its token until the graph is freed;
    backward through it still runs after close()."""
    torch = pytest.importorskip("torch")
    from ct_forward_projection_amd import autograd as A
    proj = fake_projector()
    rho = torch.ones(2, 3, requires_grad=True)
    y = A._keep_function().apply(rho, proj, proj._graph_token())
    loss = (3.0 * y).sum()
    proj.close()
    assert proj.destroyed == 0 and proj._close_pending
    loss.backward()
    assert torch.equal(rho.grad, torch.full((2, 3), 3.0))
    del y, loss
    gc.collect()
    assert proj.destroyed == 1
