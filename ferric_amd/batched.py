"""Batched local-MP2 pair domains on the device (one persistent launch over all of them).

Each domain is an independent closed-shell RI-MP2 pair-energy problem: its B
(row-major (naux_d, nocc_d*nvir_d)) lies at ``b_offset`` of one flat FP64 buffer
and its orbital energies [occupied..., virtual...] at ``eps_offset`` of another.
The reference walks its pair domains serially on the host
(crates/ferric-mp2/src/dlpno_mp2.rs:718-745); here they share one persistent
launch that pulls work items from per-XCD queues. ``use_graph`` captures the launch sequence in a hipGraph on first
use for a descriptor set and replays it afterwards.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import LayoutError, Precision, check, lib


_DOMAIN_DTYPE = np.dtype([("b_offset", "<u8"), ("eps_offset", "<u8"), ("nocc", "<u4"),
                          ("nvir", "<u4"), ("naux", "<u4"), ("reserved", "<u4")])


def descriptors(nocc, nvir, naux, b_offset, eps_offset) -> "ctypes.Array":
    """The fh_domain array (built as one numpy record array, then viewed as ctypes)."""
    n = len(nocc)
    if _DOMAIN_DTYPE.itemsize != ctypes.sizeof(_lib.DomainC):
        raise LayoutError("fh_domain layout differs from the binding's record")
    rec = np.zeros(n, dtype=_DOMAIN_DTYPE)
    for name, v in (("b_offset", b_offset), ("eps_offset", eps_offset), ("nocc", nocc),
                    ("nvir", nvir), ("naux", naux)):
        v = np.asarray(v)
        if v.shape != (n,) or (v.size and int(v.min()) < 0):
            raise LayoutError(f"domain field {name} must be {n} non-negative integers")
        if name in ("nocc", "nvir", "naux") and v.size and int(v.max()) >= 1 << 32:
            raise LayoutError(f"domain field {name} beyond 32 bits")
        rec[name] = v
    return (_lib.DomainC * n).from_buffer_copy(rec.tobytes())


def domain_energies(b_flat, eps_flat, nocc, nvir, naux, b_offset, eps_offset,
                    precision: Precision = Precision.FP64, use_graph: bool = True):
    """(e_os, e_ss) arrays, one entry per domain.

    ``b_flat`` and ``eps_flat`` are FP64 torch tensors on the device (or numpy
    arrays, copied to the device for the call)."""
    import torch

    if not (hasattr(b_flat, "is_cuda") and b_flat.is_cuda):
        b_flat = torch.from_numpy(np.ascontiguousarray(b_flat, dtype=np.float64)).cuda()
    if not (hasattr(eps_flat, "is_cuda") and eps_flat.is_cuda):
        eps_flat = torch.from_numpy(np.ascontiguousarray(eps_flat, dtype=np.float64)).cuda()
    if b_flat.dtype != torch.float64 or eps_flat.dtype != torch.float64:
        raise LayoutError("batched domains take FP64 buffers")
    if not (b_flat.is_contiguous() and eps_flat.is_contiguous()):
        raise LayoutError("batched buffers must be contiguous")
    desc = descriptors(nocc, nvir, naux, b_offset, eps_offset)
    n = len(desc)
    e_os = np.zeros(n)
    e_ss = np.zeros(n)
    stream = torch.cuda.current_stream(b_flat.device).cuda_stream
    st = lib().fh_rimp2_batched_dev(
        desc, n, ctypes.c_void_p(b_flat.data_ptr()), b_flat.numel(),
        ctypes.c_void_p(eps_flat.data_ptr()), eps_flat.numel(), int(Precision(precision)),
        1 if use_graph else 0, ctypes.c_void_p(stream), e_os.ctypes.data_as(_lib._c_dp),
        e_ss.ctypes.data_as(_lib._c_dp))
    check(st, "fh_rimp2_batched_dev")
    return e_os, e_ss


def domain_set_energies(ds, precision: Precision = Precision.FP64, use_graph: bool = True,
                        b_dev=None, eps_dev=None):
    """Convenience wrapper over a :class:`ferric_amd.synthetic.DomainSet`."""
    return domain_energies(ds.b_flat if b_dev is None else b_dev,
                           ds.eps_flat if eps_dev is None else eps_dev,
                           ds.nocc, ds.nvir, ds.naux, ds.b_offset, ds.eps_offset,
                           precision, use_graph)
