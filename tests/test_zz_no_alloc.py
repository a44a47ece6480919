"""The weight gradient's entry point that allocates nothing
(cl_conv3d_bwd_weight_noalloc: partial sums in the library's file-scope
device array) against the one with a caller workspace
(cl_conv3d_bwd_weight), bit for bit, on two calls in a row, the second
smaller; and the size limit refused before any launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(rng, nb, dims):
    import torch
    nv = dims[0] * dims[1] * dims[2]
    x = torch.tensor(rng.standard_normal((nb, nv, 16)), dtype=torch.float32).to(torch.bfloat16).cuda()
    dy = torch.tensor(rng.standard_normal((nb, nv, 16)), dtype=torch.float32).to(torch.bfloat16).cuda()
    return x, dy


def test_noalloc_weight_gradient_matches_workspace_variant():
    import torch
    from citcom_amd import learn
    L, P = learn._cl(), learn._p
    assert L.cl_scratch_bytes() == 16 << 20
    rng = np.random.default_rng(11)
    g_ref = torch.empty(16 * 27 * 16 + 16, dtype=torch.float32, device="cuda")
    g_new = torch.empty_like(g_ref)
    # the second call is smaller: the array keeps the first call's partial
    # sums beyond the second call's chunks, which must not leak into its sum
    for nb, dims in ((6, (17, 13, 11)), (2, (5, 6, 7))):
        print(f"[gpu-test] noalloc weight gradient: nb {nb} dims {dims}", flush=True)
        x, dy = _inputs(rng, nb, dims)
        ws = torch.empty(L.cl_conv3d_bwd_weight_workspace(nb, *dims) // 4, dtype=torch.float32, device="cuda")
        assert L.cl_conv3d_bwd_weight_workspace(nb, *dims) <= L.cl_scratch_bytes()
        assert L.cl_conv3d_bwd_weight(P(x), P(dy), nb, *dims, P(g_ref), P(ws), None) == 0
        g_new.fill_(float("nan"))
        assert L.cl_conv3d_bwd_weight_noalloc(P(x), P(dy), nb, *dims, P(g_new), None) == 0, L.cl_last_error()
        torch.cuda.synchronize()
        a, b = g_ref.cpu().numpy(), g_new.cpu().numpy()
        assert np.all(np.isfinite(b))
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_noalloc_refuses_oversize():
    import torch
    from citcom_amd import learn
    L, P = learn._cl(), learn._p
    dims = (65, 65, 65)
    nb = 4   # 4 * ceil(274625 / 256) chunks of 27.7 KB: far above 16 MiB
    assert L.cl_conv3d_bwd_weight_workspace(nb, *dims) > L.cl_scratch_bytes()
    x = torch.zeros((1, 16), dtype=torch.bfloat16, device="cuda")   # never read: refused before launch
    g = torch.zeros(16 * 27 * 16 + 16, dtype=torch.float32, device="cuda")
    assert L.cl_conv3d_bwd_weight_noalloc(P(x), P(x), nb, *dims, P(g), None) == -1
    assert b"scratch" in L.cl_last_error()
    torch.cuda.synchronize()
    assert not g.any()
