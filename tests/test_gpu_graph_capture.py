"""sdfx_query_sdf_noalloc recorded into a HIP graph (torch.cuda.graph's stream
capture) and replayed. Every replay gives the eager call's bits, for every
acceleration method, with eager calls and synchronisations between the
replays; those eager calls rewrite the scratch the replays reuse, so each
replay is checked against a scratch another call left behind."""
import ctypes

import numpy as np
import pytest

from mesh_to_sdf_amd import synthetic as S

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def _call(sdf, acc, dq, dout, stream):
    n = ctypes.c_uint64(0)
    st = sdf._lib.lib().sdfx_query_sdf_noalloc(acc._h, dq.data_ptr(), dq.shape[0], dout.data_ptr(), ctypes.byref(n),
                                               ctypes.c_void_p(stream.cuda_stream))
    return st, int(n.value)


def _queries(v, f, n, seed):
    return np.ascontiguousarray(np.concatenate([
        S.uniform_queries(v, n - n // 2, seed),
        S.surface_queries(v, f, n // 2, 0.01 * S.bbox_diagonal(v), seed + 10)]), dtype=np.float32)


@pytest.mark.parametrize("method", range(6))
def test_graph_replays_equal_eager(sdf, method):
    import torch
    m = sdf.ALL_SDF_ACCELERATION_METHODS[method]
    v, f = S.noise_sphere(4, 0.05)
    acc = sdf.build_sdf_acceleration(v, sdf.Topology.TriangleList(f), m)
    n = 60_000 if method != 0 else 20_000
    q, q2 = _queries(v, f, n, 21), _queries(v, f, n // 3, 22)
    ref, ref2 = sdf.query_sdf(acc, q), sdf.query_sdf(acc, q2)
    dq, dq2 = torch.from_numpy(q).cuda(), torch.from_numpy(q2).cuda()
    dout = torch.full((len(q),), float("nan"), dtype=torch.float32, device="cuda")
    dout2 = torch.full((len(q2),), float("nan"), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    # the eager call (it also resolves the scratch array's address, which is not a capturable call)
    st, cnt = _call(sdf, acc, dq, dout, torch.cuda.current_stream())
    torch.cuda.synchronize()
    assert st == 0 and cnt == n, sdf._lib.lib().sdfx_last_error()
    assert np.array_equal(bits(dout.cpu().numpy()), bits(ref))

    g = torch.cuda.CUDAGraph()
    dout.fill_(float("nan"))
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        st, cnt = _call(sdf, acc, dq, dout, torch.cuda.current_stream())
    assert st == 0 and cnt == n, sdf._lib.lib().sdfx_last_error()
    torch.cuda.synchronize()
    assert torch.isnan(dout).all().item(), "capture must not run the work"
    for rep in range(3):
        dout.fill_(float("nan"))
        dout2.fill_(float("nan"))
        st, cnt = _call(sdf, acc, dq2, dout2, torch.cuda.current_stream())  # eager, on the same scratch
        torch.cuda.synchronize()
        assert st == 0 and cnt == len(q2)
        assert np.array_equal(bits(dout2.cpu().numpy()), bits(ref2)), rep
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(bits(dout.cpu().numpy()), bits(ref)), rep
