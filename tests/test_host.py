"""CPU: host plan, module buffers, argument validation, and the C-ABI surface.

No GPU compute happens here: library calls below only exercise the argument
checks that run before anything is enqueued.
"""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

import inputs as gi
import torchkbnufft_amd as tkbn
from conftest import ROOT, golden
from torchkbnufft_amd import _lib
from torchkbnufft_amd._nufft import utils


@pytest.mark.parametrize("idx", range(len(gi.PLAN_CASES)))
def test_plan_buffers_bit_identical_to_reference(idx):
    im, grid, J, L, w = gi.PLAN_CASES[idx]
    p = golden("plans_c128.npz")
    out = utils.init_fn(im, grid, J, None, L, w, 0.0, torch.complex128)
    tables = out[0]
    for k in range(len(im)):
        key = f"p{idx}_table{k}"
        if key in p.files:
            assert np.array_equal(tables[k].numpy(), p[key])
        else:
            j = int(p[key + "_same_as"])
            assert torch.equal(tables[k], tables[j])
    sc = utils.compute_scaling_coefs(out[1].tolist(), out[2].tolist(), out[4].tolist(), out[8].tolist(), out[7].tolist())
    d = len(im)
    for k in range(d):
        sl = [0] * d
        sl[k] = slice(None)
        assert np.array_equal(sc.numpy()[tuple(sl)], p[f"p{idx}_scaling_line{k}"])
    if f"p{idx}_scaling_coef" in p.files:
        assert np.array_equal(sc.numpy(), p[f"p{idx}_scaling_coef"])


def test_module_buffers_match_reference_layout():
    ob = tkbn.KbNufft(im_size=(12, 10), numpoints=(4, 6), dtype=torch.complex64)
    sd = ob.state_dict()
    assert set(sd) == {
        "table_0", "table_1", "im_size", "grid_size", "n_shift", "numpoints", "offsets",
        "table_oversamp", "order", "alpha", "scaling_coef",
    }
    assert sd["table_0"].shape == (4 * 1024 + 1,) and sd["table_1"].shape == (6 * 1024 + 1,)
    assert sd["table_0"].dtype == torch.complex64 and sd["scaling_coef"].dtype == torch.complex64
    assert sd["n_shift"].dtype == torch.float32 and sd["alpha"].dtype == torch.float32
    for k in ("im_size", "grid_size", "numpoints", "offsets", "table_oversamp"):
        assert sd[k].dtype == torch.int64
    assert sd["offsets"].shape == (24, 2)
    # itertools.product order: last dimension fastest
    assert sd["offsets"][:7].tolist() == [[0, 0], [0, 1], [0, 2], [0, 3], [0, 4], [0, 5], [1, 0]]
    assert sd["grid_size"].tolist() == [24, 20] and sd["n_shift"].tolist() == [6, 5]
    ob2 = tkbn.KbNufft(im_size=(12, 10), numpoints=(4, 6), dtype=torch.complex64)
    ob2.load_state_dict(sd)


def test_module_to_maps_complex_and_real():
    ob = tkbn.KbInterp(im_size=(8, 8), dtype=torch.complex64).to(torch.complex128)
    assert ob.table_0.dtype == torch.complex128 and ob.n_shift.dtype == torch.float64
    assert ob.offsets.dtype == torch.int64
    ob = ob.to(torch.float32)
    assert ob.table_0.dtype == torch.complex64 and ob.alpha.dtype == torch.float32
    with pytest.raises(TypeError):
        ob.to(torch.int32)


def _header_functions():
    with open(os.path.join(ROOT, "include", "tkbn_hip.h")) as f:
        text = f.read()
    return re.findall(r"TKBN_API\s+[\w\s\*]+?\b(tkbn_\w+)\s*\(", text)


def test_library_exports_every_header_symbol():
    names = _header_functions()
    assert len(names) >= 8
    lib = _lib.load()
    for n in names:
        assert hasattr(lib, n), n
    assert lib.tkbn_abi_version() == _lib.ABI_VERSION
    assert set(names) <= set(_lib._SIGNATURES)


def _desc(**kw):
    d = _lib.InterpDesc()
    d.dtype = kw.get("dtype", _lib.TKBN_F32)
    d.ndim = kw.get("ndim", 2)
    for i in range(3):
        d.grid_size[i] = 8
        d.numpoints[i] = kw.get("J", 6)
        d.table_oversamp[i] = 1024
        d.tables[i] = 1  # never dereferenced: validation fails first or sizes are empty
    d.batch, d.coils, d.klength = kw.get("B", 1), 1, kw.get("M", 4)
    return d


def test_abi_rejects_bad_descriptors_without_gpu_work():
    lib = _lib.load()
    assert lib.tkbn_interp_fwd(None, None, None, None, None, 0, None) == _lib.TKBN_ERR_INVALID
    assert b"NULL" in lib.tkbn_last_error()
    assert lib.tkbn_interp_fwd(ctypes.byref(_desc(ndim=4)), 1, 1, 1, None, 0, None) == _lib.TKBN_ERR_UNSUPPORTED
    assert lib.tkbn_interp_fwd(ctypes.byref(_desc(J=17)), 1, 1, 1, None, 0, None) == _lib.TKBN_ERR_UNSUPPORTED
    assert lib.tkbn_interp_fwd(ctypes.byref(_desc(dtype=5)), 1, 1, 1, None, 0, None) == _lib.TKBN_ERR_INVALID
    assert lib.tkbn_interp_fwd(ctypes.byref(_desc(B=-1)), 1, 1, 1, None, 0, None) == _lib.TKBN_ERR_INVALID
    # empty batch is a no-op
    assert lib.tkbn_interp_fwd(ctypes.byref(_desc(B=0)), 1, 1, 1, None, 0, None) == _lib.TKBN_OK
    im = _lib.i64_array([9, 9])
    gs = _lib.i64_array([8, 8])
    st = lib.tkbn_apod_pad(0, 2, im, gs, 1, 1, 1, 1, None, 1, 1.0, 1, None)
    assert st == _lib.TKBN_ERR_INVALID and b"im_size" in lib.tkbn_last_error()
    st = lib.tkbn_apod_pad(0, 2, gs, gs, 1, 2, 1, 1, None, 1, 1.0, 1, None)
    assert st == _lib.TKBN_ERR_INVALID  # image_coils != coils without smaps
    assert lib.tkbn_crop_deapod(0, 0, gs, gs, 1, 1, 1, None, 1, 1.0, 1, None) == _lib.TKBN_ERR_UNSUPPORTED
    with pytest.raises(NotImplementedError):
        _lib.check(_lib.TKBN_ERR_UNSUPPORTED, "x")
    with pytest.raises(ValueError):
        _lib.check(_lib.TKBN_ERR_INVALID, "x")


def test_python_validation_matches_reference_errors():
    ob = tkbn.KbInterp(im_size=(8, 8))
    image = torch.randn(1, 1, 8, 8, dtype=torch.complex64)
    with pytest.raises(ValueError, match="omega must have 2 or 3 dimensions"):
        ob(image, torch.zeros(1, 1, 2, 5))
    with pytest.raises(ValueError, match="last dimension must be size 2"):
        ob(torch.randn(1, 1, 8, 8, 3), torch.zeros(2, 5))
    with pytest.raises(ValueError, match="batch dimension must match"):
        ob(image, torch.zeros(3, 2, 5))
    nu = tkbn.KbNufft(im_size=(8, 8))
    with pytest.raises(TypeError, match="smaps"):
        nu(image, torch.zeros(2, 5), smaps=torch.randn(1, 2, 8, 8, dtype=torch.complex128))
    with pytest.raises(TypeError, match="2-tuple"):
        tkbn.KbInterpAdjoint(im_size=(8, 8))(
            torch.randn(1, 1, 5, dtype=torch.complex64), torch.zeros(2, 5), interp_mats=torch.zeros(1)
        )


def test_cpu_tensors_are_refused_not_emulated():
    ob = tkbn.KbInterp(im_size=(8, 8))
    with pytest.raises(RuntimeError, match="GPU"):
        ob(torch.randn(1, 1, 16, 16, dtype=torch.complex64), torch.zeros(2, 5))
    adj = tkbn.KbNufftAdjoint(im_size=(8, 8))
    with pytest.raises(RuntimeError, match="GPU"):
        adj(torch.randn(1, 1, 5, dtype=torch.complex64), torch.zeros(2, 5))


def test_dcomp_and_toeplitz_validation_matches_reference_errors():
    with pytest.raises(ValueError, match="ktraj must have 2 or 3 dimensions"):
        tkbn.calc_density_compensation_function(torch.zeros(5), (8, 8))
    with pytest.raises(ValueError, match="Unrecognized k-space shape"):
        tkbn.calc_toeplitz_kernel(torch.zeros(5), (8, 8))
    with pytest.raises(ValueError, match="Unrecognized weights dimension"):
        tkbn.calc_toeplitz_kernel(torch.zeros(2, 5), (8, 8), weights=torch.ones(5))
    with pytest.raises(ValueError, match="same batch size"):
        tkbn.calc_toeplitz_kernel(torch.zeros(3, 2, 5), (8, 8), weights=torch.ones(2, 5))
    toep = tkbn.ToepNufft()
    image = torch.randn(1, 1, 8, 8, dtype=torch.complex64)
    with pytest.raises(TypeError, match="same dtype"):
        toep(image, torch.randn(16, 16, dtype=torch.complex128))
    with pytest.raises(TypeError, match="smaps"):
        toep(image, torch.randn(16, 16, dtype=torch.complex64), smaps=torch.randn(1, 2, 8, 8, dtype=torch.complex128))
    with pytest.raises(ValueError, match="last dimension must be size 2"):
        toep(torch.randn(1, 1, 8, 8, 3), torch.randn(16, 16, 3))
    with pytest.raises(ValueError, match="same batch size as image"):
        toep(image, torch.randn(3, 16, 16, dtype=torch.complex64))
    with pytest.raises(ValueError, match="Only option for norm is 'ortho'"):
        tkbn.functional.fft_filter(image, torch.randn(16, 16, dtype=torch.complex64), norm="forward")
    with pytest.raises(RuntimeError, match="GPU"):
        toep(image, torch.randn(16, 16, dtype=torch.complex64))


def test_math_helpers():
    a = torch.randn(3, 4, dtype=torch.complex128)
    b = torch.randn(3, 4, dtype=torch.complex128)
    # the reference returns sum(val2 * conj(val1)) (_math.py:160, conj_complex_mult(val2, val1))
    assert torch.allclose(tkbn.inner_product(a, b), torch.sum(b * a.conj()))
    e0 = torch.zeros(2, dtype=torch.complex128)
    e0[0] = 1.0
    assert complex(tkbn.inner_product(e0 * 1j, e0 * 2.0)) == pytest.approx(-2j)  # conj on the first argument
    ar, br = torch.view_as_real(a), torch.view_as_real(b)
    ip = tkbn.inner_product(ar, br)
    assert torch.allclose(torch.complex(ip[0], ip[1]), torch.sum(b * a.conj()))
    assert torch.allclose(torch.view_as_complex(tkbn.complex_mult(ar, br)), a * b)
    assert torch.allclose(torch.view_as_complex(tkbn.conj_complex_mult(ar, br)), a * b.conj())
    assert torch.allclose(tkbn.imag_exp(torch.tensor([0.5])), torch.exp(1j * torch.tensor([0.5])))
    # imag_exp's real form stacks (cos, sin) last whatever dim is (_math.py:133)
    assert tkbn.imag_exp(torch.zeros(3, 5), dim=0, return_complex=False).shape == (3, 5, 2)


def test_complex_sign_matches_atan2_definition():
    """complex_sign = exp(i atan2(im, re)) (_math.py:60-82): 1 at 0, -1 at -0.0, unit elsewhere."""
    v = torch.tensor([0.0 + 0.0j, 3.0 + 4.0j, -2.0 + 0.0j, 0.0 - 5.0j], dtype=torch.complex128)
    s = tkbn.complex_sign(v)
    assert torch.allclose(s, torch.tensor([1.0, 0.6 + 0.8j, -1.0, -1j], dtype=torch.complex128))
    neg0 = torch.view_as_complex(torch.tensor([[-0.0, 0.0]], dtype=torch.float64))
    assert torch.allclose(tkbn.complex_sign(neg0), torch.tensor([-1.0 + 0j], dtype=torch.complex128))
    r = torch.view_as_real(v)  # real input, (re, im) along the last dim
    assert torch.allclose(torch.view_as_complex(tkbn.complex_sign(r).contiguous()), s)
    with pytest.raises(ValueError):
        tkbn.complex_sign(torch.zeros(4, 3))


@pytest.mark.parametrize("rows,ndim,mode", [(1, 2, 0), (64, 2, 0), (64, 2, 1), (3, 3, 1), (1, 1, 0)])
def test_plan_reuse_refused_for_unbuilt_workspace(rows, ndim, mode):
    """TKBN_PLAN_REUSE on a workspace that holds no plan for this geometry is refused
    with TKBN_ERR_INVALID before anything is enqueued (the pointer is never touched)."""
    lib = _lib.load()
    d = _desc(ndim=ndim, M=5000)
    d.coils = rows
    for i in range(3):
        d.grid_size[i] = 640 if ndim == 2 else 64
    d.flags = _lib.TKBN_PLAN_REUSE
    fake_ws = 1 << 40  # 16-byte aligned, never dereferenced
    big = 1 << 40
    assert lib.tkbn_interp_adj(ctypes.byref(d), 1, 1, 1, mode, fake_ws, big, None) == _lib.TKBN_ERR_INVALID
    assert b"TKBN_PLAN_REUSE" in lib.tkbn_last_error()
    if lib.tkbn_interp_fwd_workspace_bytes(ctypes.byref(d)) > 0:
        assert lib.tkbn_interp_fwd(ctypes.byref(d), 1, 1, 1, fake_ws, big, None) == _lib.TKBN_ERR_INVALID
        assert b"TKBN_PLAN_REUSE" in lib.tkbn_last_error()


@pytest.mark.parametrize("ndim", [1, 2, 3])
@pytest.mark.parametrize("rows", [1, 8, 64])
def test_workspace_size_queries_are_host_only_and_positive(ndim, rows):
    """Size queries run host arithmetic only (no GPU): every geometry gives a size."""
    lib = _lib.load()
    d = _desc(ndim=ndim)
    d.coils, d.klength = rows, 1000
    for i in range(3):
        d.grid_size[i] = 64 if ndim < 3 else 24
    assert lib.tkbn_interp_adj_workspace_bytes(ctypes.byref(d), _lib.TKBN_ADJ_ATOMIC) > _lib.PLAN_CACHE_HEADER_BYTES
    assert lib.tkbn_interp_adj_workspace_bytes(ctypes.byref(d), _lib.TKBN_ADJ_DETERMINISTIC) > 0
    assert lib.tkbn_interp_fwd_workspace_bytes(ctypes.byref(d)) >= 0
    d.flags = 2
    assert lib.tkbn_interp_adj_workspace_bytes(ctypes.byref(d), 0) == -1  # unknown flag


def test_noalloc_adjoint_refuses_oversized_and_flagged_calls_on_the_host():
    """tkbn_interp_adj_noalloc rejects a problem whose workspace exceeds its static
    scratch, and any flags, before touching the GPU (pointers are never dereferenced)."""
    lib = _lib.load()
    cap = lib.tkbn_interp_adj_noalloc_max_bytes()
    assert cap == 128 << 20
    d = _desc(ndim=3)
    for i in range(3):
        d.grid_size[i] = 320
    d.coils, d.klength = 8, 1 << 24
    for mode in (_lib.TKBN_ADJ_ATOMIC, _lib.TKBN_ADJ_DETERMINISTIC):
        assert lib.tkbn_interp_adj_workspace_bytes(ctypes.byref(d), mode) > cap
        assert lib.tkbn_interp_adj_noalloc(ctypes.byref(d), 1, 1, 1, mode, None) == _lib.TKBN_ERR_WORKSPACE
        assert b"static scratch" in lib.tkbn_last_error()
    d.klength = 1000
    d.flags = _lib.TKBN_PLAN_REUSE
    assert lib.tkbn_interp_adj_noalloc(ctypes.byref(d), 1, 1, 1, 0, None) == _lib.TKBN_ERR_INVALID
    assert b"flags" in lib.tkbn_last_error()
    assert lib.tkbn_interp_adj_noalloc(None, 1, 1, 1, 0, None) == _lib.TKBN_ERR_INVALID


def test_tuning_fields_are_validated_on_the_host():
    """tile / work_target overrides: an LDS-overflowing tile or a tiny target is refused
    by the size queries (-1); a valid smaller tile gives a larger plan (more tiles)."""
    lib = _lib.load()
    d = _desc(ndim=2)
    d.coils, d.klength = 4, 5000
    for i in range(2):
        d.grid_size[i] = 256
    base = lib.tkbn_interp_adj_workspace_bytes(ctypes.byref(d), _lib.TKBN_ADJ_ATOMIC)
    assert base > 0
    d.tile[0] = d.tile[1] = 8
    small = lib.tkbn_interp_adj_workspace_bytes(ctypes.byref(d), _lib.TKBN_ADJ_ATOMIC)
    assert small > base
    d.tile[0] = d.tile[1] = 64  # (64 + 5)^2 * 16 bytes > 60 KiB
    assert lib.tkbn_interp_adj_workspace_bytes(ctypes.byref(d), _lib.TKBN_ADJ_ATOMIC) == -1
    assert lib.tkbn_interp_fwd_workspace_bytes(ctypes.byref(d)) == -1
    assert b"LDS" in lib.tkbn_last_error()
    d.tile[0] = d.tile[1] = 0
    d.work_target = 8
    assert lib.tkbn_interp_adj_workspace_bytes(ctypes.byref(d), _lib.TKBN_ADJ_ATOMIC) == -1
    d.work_target = 128  # smaller pieces: more work items in the plan
    assert lib.tkbn_interp_adj_workspace_bytes(ctypes.byref(d), _lib.TKBN_ADJ_ATOMIC) > base


def _batch_ptrs(n, base=1 << 40):
    arr = lambda k: (ctypes.c_void_p * n)(*[base + (k * n + p) * 4096 for p in range(n)])  # noqa: E731
    return [arr(k) for k in range(7)]


def test_batch_create_validates_problems_on_the_host():
    """tkbn_batch_create only does host arithmetic (pointers are recorded, never read):
    it refuses NULL pointers, multi-coil problems, mixed dimensions and forward / adjoint
    descriptors that differ in more than tuning; a batch call whose plans were never
    built is refused before anything is enqueued."""
    lib = _lib.load()
    n = 3
    descs = (_lib.InterpDesc * n)()
    for p in range(n):
        d = _desc(ndim=2, M=3000 + 500 * p)
        d.coils = 8
        for i in range(2):
            d.grid_size[i] = 128 + 64 * p
        descs[p] = d
    out = ctypes.c_void_p()
    ptrs = _batch_ptrs(n)
    assert lib.tkbn_batch_create(descs, None, n, 0, *ptrs, ctypes.byref(out)) == _lib.TKBN_OK
    assert lib.tkbn_batch_table_bytes(out) > 0 and lib.tkbn_batch_table_bytes(out) % n == 0
    assert lib.tkbn_batch_workspace_bytes(out) == 2 * (n + 1) * 4
    host = (ctypes.c_ubyte * lib.tkbn_batch_table_bytes(out))()
    assert lib.tkbn_batch_write_table(out, host) == _lib.TKBN_OK
    ws = 1 << 41
    assert lib.tkbn_batch_interp_fwd(out, 1 << 42, ws, 4096, None) == _lib.TKBN_ERR_INVALID
    assert b"TKBN_PLAN_REUSE" in lib.tkbn_last_error()
    assert lib.tkbn_batch_interp_adj(out, 1 << 42, ws, 4, None) == _lib.TKBN_ERR_WORKSPACE
    assert lib.tkbn_batch_destroy(out) == _lib.TKBN_OK
    # adjoint descriptors that differ only in tuning are accepted
    adescs = (_lib.InterpDesc * n)()
    for p in range(n):
        adescs[p] = descs[p]
        adescs[p].tile[0] = adescs[p].tile[1] = 16
        adescs[p].work_target = 256
    assert lib.tkbn_batch_create(descs, adescs, n, 0, *ptrs, ctypes.byref(out)) == _lib.TKBN_OK
    lib.tkbn_batch_destroy(out)
    adescs[1].klength += 1
    assert lib.tkbn_batch_create(descs, adescs, n, 0, *ptrs, ctypes.byref(out)) == _lib.TKBN_ERR_INVALID
    assert b"differ" in lib.tkbn_last_error()
    # a NULL pointer, a 32-row problem, a 3-D problem among 2-D ones
    bad = _batch_ptrs(n)
    bad[3][1] = None
    assert lib.tkbn_batch_create(descs, None, n, 0, *bad, ctypes.byref(out)) == _lib.TKBN_ERR_INVALID
    descs[2].coils = 32
    assert lib.tkbn_batch_create(descs, None, n, 0, *ptrs, ctypes.byref(out)) == _lib.TKBN_ERR_UNSUPPORTED
    assert b"problem 2" in lib.tkbn_last_error()
    descs[2].coils = 8
    descs[1].ndim = 3
    assert lib.tkbn_batch_create(descs, None, n, 0, *ptrs, ctypes.byref(out)) == _lib.TKBN_ERR_UNSUPPORTED
    assert lib.tkbn_batch_create(descs, None, 0, 0, *ptrs, ctypes.byref(out)) == _lib.TKBN_ERR_INVALID
