"""The dressing / MO-transform stage upstream of the pair kernels, on the device.

Mirrors ``stream_dressed_mo_band`` (crates/ferric-mp2/src/rimp2.rs:648-756):
``B[P, l*nright + r] = sum_Q M[P, Q] (C_left^T (Q|mu nu) C_right)[l, r]`` over an
optional output band of aux rows, with M the dressing matrix (the reference's
``v_inv_sqrt``). Inputs and the result are CUDA FP64 tensors, so B goes straight
into :func:`ferric_amd.rimp2.spin_components` without a host round trip.
"""
from __future__ import annotations

import ctypes

from . import _lib
from ._lib import LayoutError, check, lib


def stream_dressed_mo_band(j3, v_inv_sqrt, c_left, c_right, output_band=None):
    """(band, nleft*nright) device tensor of the dressed MO-pair factor.

    j3: (naux, nao, nao) three-centre integrals (Q|mu nu), symmetric in mu, nu;
    v_inv_sqrt: (naux, naux) dressing matrix; c_left (nao, nleft), c_right
    (nao, nright) MO coefficients; output_band = (p0, p1) keeps rows p0..p1-1.
    """
    import torch

    for name, t in (("j3", j3), ("v_inv_sqrt", v_inv_sqrt), ("c_left", c_left),
                    ("c_right", c_right)):
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64):
            raise LayoutError(f"{name} must be a CUDA float64 tensor")
    if j3.dim() != 3 or j3.shape[1] != j3.shape[2]:
        raise LayoutError("j3 must be (naux, nao, nao)")
    naux, nao = j3.shape[0], j3.shape[1]
    if c_left.shape[0] != nao or c_right.shape[0] != nao:
        raise LayoutError("MO coefficients must have nao rows")
    if tuple(v_inv_sqrt.shape) != (naux, naux):
        raise LayoutError("v_inv_sqrt must be (naux, naux)")
    p0, p1 = output_band if output_band is not None else (0, naux)
    if not 0 <= p0 <= p1 <= naux:
        raise LayoutError(f"output band {output_band} outside [0, {naux}]")
    j3, cl, cr = j3.contiguous(), c_left.contiguous(), c_right.contiguous()
    m = v_inv_sqrt[p0:p1].contiguous()
    nleft, nright = cl.shape[1], cr.shape[1]
    out = torch.empty((p1 - p0, nleft * nright), dtype=torch.float64, device=j3.device)
    stream = torch.cuda.current_stream(j3.device).cuda_stream
    st = lib().fh_dress_b_ov_dev(ctypes.c_void_p(j3.data_ptr()), naux, nao,
                                 ctypes.c_void_p(cl.data_ptr()), nleft,
                                 ctypes.c_void_p(cr.data_ptr()), nright,
                                 ctypes.c_void_p(m.data_ptr()), p1 - p0, ctypes.c_void_p(stream),
                                 ctypes.c_void_p(out.data_ptr()))
    check(st, "fh_dress_b_ov_dev")
    return out
