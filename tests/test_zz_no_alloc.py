"""tkbn_interp_adj_noalloc (static device scratch, no allocation) against tkbn_interp_adj
(caller workspace), called through the C ABI: two calls in a row, the second on a
smaller problem, in both adjoint modes, plus the oracle."""
import ctypes

import pytest
import torch

import torchkbnufft_amd as tkbn
import workloads as W
from conftest import rel_l2
from oracle import table_path as tp
from torchkbnufft_amd import _lib
from torchkbnufft_amd._nufft import interp as _interp

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")


def _problem(im, nspokes, nread, rows, seed):
    gen = torch.Generator().manual_seed(seed)
    om = W.radial_golden(nspokes, nread)
    y = W.cplx((1, rows, om.shape[1]), gen)
    return om, y


def _adjoint_both(im, om, y, mode):
    """(noalloc output, tkbn_interp_adj output) for the same inputs."""
    lib = _lib.load()
    ad = tkbn.KbInterpAdjoint(im_size=im, dtype=torch.complex64).to(DEV)
    tabs = [t.contiguous() for t in ad._tables()]
    grid = tuple(2 * n for n in im)
    yd, od = y.to(DEV).contiguous(), om.to(DEV).contiguous()
    desc = _interp.make_desc(yd, od, tabs, ad.n_shift, ad.numpoints, ad.table_oversamp, grid, y.shape[0], y.shape[1])
    out_na = torch.full((y.shape[0], y.shape[1]) + grid, float("nan"), dtype=torch.complex64, device=DEV)
    out_ws = torch.empty_like(out_na)
    need = lib.tkbn_interp_adj_workspace_bytes(ctypes.byref(desc), mode)
    assert 0 < need <= lib.tkbn_interp_adj_noalloc_max_bytes()
    ws = torch.zeros(need, dtype=torch.uint8, device=DEV)
    stream = torch.cuda.current_stream().cuda_stream
    with torch.cuda.device(DEV):
        st = lib.tkbn_interp_adj_noalloc(ctypes.byref(desc), yd.data_ptr(), od.data_ptr(), out_na.data_ptr(), mode,
                                         stream)
        _lib.check(st, "tkbn_interp_adj_noalloc")
        st = lib.tkbn_interp_adj(ctypes.byref(desc), yd.data_ptr(), od.data_ptr(), out_ws.data_ptr(), mode,
                                 ws.data_ptr(), ws.numel(), stream)
        _lib.check(st, "tkbn_interp_adj")
    torch.cuda.synchronize()
    return out_na, out_ws


@pytest.mark.parametrize("mode", [_lib.TKBN_ADJ_ATOMIC, _lib.TKBN_ADJ_DETERMINISTIC])
def test_noalloc_adjoint_matches_workspace_adjoint(mode):
    cases = [((160, 160), 100, 320, 40, 71), ((48, 48), 24, 96, 3, 72)]  # second call: smaller
    for im, nsp, nrd, rows, seed in cases:
        print(f"[case] noalloc adjoint mode={mode} im={im} samples={nsp * nrd} rows={rows}", flush=True)
        om, y = _problem(im, nsp, nrd, rows, seed)
        got, want = _adjoint_both(im, om, y, mode)
        assert torch.isfinite(torch.view_as_real(got)).all()
        if mode == _lib.TKBN_ADJ_DETERMINISTIC:
            assert torch.equal(got, want)
        else:
            assert rel_l2(got, want) < 1e-6
        plan = tp.Plan(im, dtype=torch.complex64)
        sel = [0, rows - 1]
        err = rel_l2(got[:, sel], tp.kb_table_interp_adjoint(plan, y[:, sel], om))
        print(f"[noalloc] vs workspace entry ok, vs oracle {err:.2e}", flush=True)
        assert err < 1e-5
