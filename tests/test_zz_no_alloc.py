"""ngmi_mlp_forward_noalloc (the guess network's forward pass on the
library's fixed per-device scratch, for callers that cannot allocate)
against ngmi_mlp_forward through the C ABI: bit-identical outputs on two
calls in a row, the second with fewer rows than the first, and a clear
error above the scratch maximum."""
import ctypes

import numpy as np
import pytest

from engine_path import ngmi

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu


def _noalloc(net, X, y, M):
    L = ngmi.lib()
    vp = ctypes.c_void_p
    L.ngmi_mlp_forward_noalloc.argtypes = [vp, vp, vp, ctypes.c_int, vp]
    return L.ngmi_mlp_forward_noalloc(net.h, X, y, M, None)


def test_noalloc_forward_matches_forward():
    print('case: guess network forward, no-alloc entry vs ngmi_mlp_forward, M = 12345 then 700',
          flush=True)
    net = ngmi.GuessNet(seed=7)
    g = np.random.default_rng(3)
    for M in (12345, 700):
        X = torch.from_numpy(g.normal(0, 1, (M, 16)).astype(np.float32)).cuda()
        y_ref = torch.full((M,), np.nan, dtype=torch.float32, device='cuda')
        y_new = torch.full((M,), np.nan, dtype=torch.float32, device='cuda')
        net.forward(X.data_ptr(), y_ref.data_ptr(), M)
        assert _noalloc(net, X.data_ptr(), y_new.data_ptr(), M) == 0, ngmi.last_error()
        torch.cuda.synchronize()
        a, b = y_ref.cpu().numpy(), y_new.cpu().numpy()
        print('M %d: max |y_noalloc - y_forward| %.3e, bit-equal %s' % (M, np.abs(a - b).max(),
                                                                        np.array_equal(a, b)), flush=True)
        assert np.isfinite(b).all()
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_noalloc_forward_rejects_too_many_rows():
    net = ngmi.GuessNet(seed=7)
    M = 16385
    X = torch.zeros((M, 16), dtype=torch.float32, device='cuda')
    y = torch.zeros(M, dtype=torch.float32, device='cuda')
    rc = _noalloc(net, X.data_ptr(), y.data_ptr(), M)
    assert rc == ngmi.E_PARMVAL if hasattr(ngmi, 'E_PARMVAL') else rc == 11
    assert 'scratch maximum' in ngmi.last_error()
