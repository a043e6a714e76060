"""Write golden output vectors from the reference torchkbnufft (run in the dev container only).

Usage (from the repository root, with the reference checkout at /root/reference):

    PYTHONPATH=/root/reference python tests/golden/make_golden.py [fixtures] [c64] [plans] [grad] [spmat]

Why the reference is run here
-----------------------------
* The reference's own known-answer fixtures (``tests/data/interp_data.pkl``,
  ``tests/data/nufft_data.pkl``) are pickles that ``torch.load(weights_only=True)``
  refuses, so they are not loaded.  The same cases (shapes, ``arange + N(0,1)``
  inputs, ``U(-pi, pi)`` trajectories; ``tests/data/create_old_data.py:17-119``)
  are regenerated and the reference's outputs written as ``.npz``.
* Nothing but the reference itself pins its complex64 float32 index arithmetic
  (division, half-to-even rounding, floor) and so its exact complex64 outputs.

Inputs come from ``inputs.py`` (seeded); only outputs are stored, as plain
arrays that ``numpy.load(allow_pickle=False)`` reads.  Nothing here runs on a
GPU machine.
"""
from __future__ import annotations

import os
import sys
import warnings

import numpy as np
import torch

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
warnings.filterwarnings("ignore")
import inputs as gi  # noqa: E402
import torchkbnufft as tkbn  # noqa: E402  (reference, from PYTHONPATH)

OUT = os.path.dirname(os.path.abspath(__file__))


def fixture_cases():
    out = {}
    torch.set_default_dtype(torch.double)
    for i in range(len(gi.INTERP_FIXTURES)):
        im_size, image, ktraj = gi.interp_fixture_inputs(i)
        out[f"interp{i}_kdata"] = tkbn.KbInterp(im_size=im_size, grid_size=im_size)(image, ktraj).numpy()
    for i in range(len(gi.NUFFT_FIXTURES)):
        im_size, image, ktraj = gi.nufft_fixture_inputs(i)
        kdata = tkbn.KbNufft(im_size=im_size)(image, ktraj)
        out[f"nufft{i}_kdata"] = kdata.numpy()
        if i != 1:  # the 78x96 adjoint is 600 KB; the other two cover 1-D and 3-D
            out[f"nufft{i}_adj"] = tkbn.KbNufftAdjoint(im_size=im_size)(kdata, ktraj).numpy()
    torch.set_default_dtype(torch.float32)
    np.savez_compressed(os.path.join(OUT, "fixtures_f64.npz"), **out)


def c64_cases():
    out = {}
    for case in gi.C64_CASES:
        name, im, grid, J, L, B, C, mk, use_smaps, norm, batched = case
        inp = gi.c64_inputs(case)
        kw = dict(im_size=im, grid_size=grid, numpoints=J, table_oversamp=L, dtype=torch.complex64)
        fw, ad = tkbn.KbNufft(**kw), tkbn.KbNufftAdjoint(**kw)
        iw, ia = tkbn.KbInterp(**kw), tkbn.KbInterpAdjoint(**kw)
        k, sm = inp["ktraj"], inp["smaps"]
        p = name + "_"
        out[p + "nufft"] = fw(inp["x"], k, smaps=sm, norm=norm).numpy()
        out[p + "nufft_adj"] = ad(inp["y_in"], k, smaps=sm, norm=norm).numpy()
        out[p + "interp"] = iw(inp["grid_in"], k).numpy()
        out[p + "interp_adj"] = ia(inp["y_in"], k).numpy()
    np.savez_compressed(os.path.join(OUT, "golden_c64.npz"), **out)


def plan_cases():
    out = {}
    for i, (im, grid, J, L, w) in enumerate(gi.PLAN_CASES):
        ob = tkbn.KbNufft(
            im_size=im, grid_size=grid, numpoints=J, table_oversamp=L, kbwidth=w, dtype=torch.complex128
        )
        d = len(im)
        p = f"p{i}_"
        for k in range(d):
            t = getattr(ob, f"table_{k}").numpy()
            same = [j for j in range(k) if np.array_equal(t, out.get(p + f"table{j}", None))]
            if same:
                out[p + f"table{k}_same_as"] = np.array(same[0])
            else:
                out[p + f"table{k}"] = t
        sc = ob.scaling_coef.numpy()
        # the reference builds scaling_coef as an outer product of 1-D factors
        # (utils.py:231-252); store the factors' lines through index 0 and the corner
        for k in range(d):
            idx = [0] * d
            idx[k] = slice(None)
            out[p + f"scaling_line{k}"] = sc[tuple(idx)]
        if d <= 2 and sc.size <= 10000:
            out[p + "scaling_coef"] = sc
    np.savez_compressed(os.path.join(OUT, "plans_c128.npz"), **out)


def grad_cases():
    im, ktraj, x, smaps, y = gi.grad_inputs()
    fw = tkbn.KbNufft(im_size=im, dtype=torch.complex64)
    x = x.clone().requires_grad_(True)
    loss = torch.sum(torch.abs(fw(x, ktraj, smaps=smaps) - y) ** 2)
    loss.backward()
    np.savez_compressed(
        os.path.join(OUT, "grad_c64.npz"), loss=np.array(loss.item()), grad_x=x.grad.numpy()
    )


def spmat_cases():
    out = {}
    for i, (im, grid, J, order, m, dtype) in enumerate(gi.SPMAT_CASES):
        omega = gi.spmat_inputs(i)
        re, imag = tkbn.calc_tensor_spmatrix(omega, im, grid_size=grid, numpoints=J, order=order)
        # coalesced (row-major sorted, duplicates summed): the stored form of the matrix
        mat = torch.sparse_coo_tensor(re._indices(), torch.complex(re._values(), imag._values()), re.shape)
        mat = mat.coalesce()
        out[f"s{i}_indices"] = mat.indices().numpy()
        out[f"s{i}_values"] = mat.values().numpy()
        out[f"s{i}_shape"] = np.array(re.shape)
        out[f"s{i}_nnz_raw"] = np.array(re._nnz())
    np.savez_compressed(os.path.join(OUT, "spmat.npz"), **out)


if __name__ == "__main__":
    # optional arguments select parts: fixtures c64 plans grad spmat (default: all)
    parts = sys.argv[1:] or ["fixtures", "c64", "plans", "grad", "spmat"]
    for name, fn in (
        ("fixtures", fixture_cases),
        ("c64", c64_cases),
        ("plans", plan_cases),
        ("grad", grad_cases),
        ("spmat", spmat_cases),
    ):
        if name in parts:
            fn()
    for f in sorted(os.listdir(OUT)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(OUT, f)))
