"""MFMA lane layouts the kernels rely on, checked on the device with one wave.

Assumed (and relied on by ferric_amd/csrc/energy*.hip):
* 16x16xK: lane l = 16 g + r supplies A[row r][k] and B[k][col r] for its k set
  (the same set for A and B), and result element v of lane l is C[4 g + v][r] (C[4 v + g][r] for
  v_mfma_f64_16x16x4f64);
* 16x16x64 INT8 (int32 result): the same row 4 g + v as the FP32-accumulating
  instructions, bytes paired between A and B lanes of one group g (exact);
* scaled FP8 (format 0 = OCP E4M3, E8M0 bias 127): bytes 0-15 of lane group g
  are k = 16g..16g+15, bytes 16-31 are k = 64+16g..; the scale of 32-wide k
  block kb of row r comes from lane 16 kb + r.
"""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROBE = os.path.join(ROOT, "tests", "hip", "build", "libmfma_probe.so")


@pytest.fixture(scope="module")
def probe():
    L = ctypes.CDLL(PROBE)
    L.probe_mfma.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                             ctypes.c_void_p, ctypes.c_void_p]
    return L


def _run(probe, kind, a, b, out, sa=None, sb=None):
    print(f"probe kind {kind}", flush=True)
    rc = probe.probe_mfma(kind, a.ctypes.data, b.ctypes.data,
                          None if sa is None else sa.ctypes.data,
                          None if sb is None else sb.ctypes.data, out.ctypes.data)
    assert rc == 0
    return out


# Row of result element v of lane 16 g + r (column r), per instruction, as the
# kernels assume it (ferric_amd/csrc/fh_device.h, out_row_*).
ROW_F64 = lambda g, v: 4 * v + g  # noqa: E731
ROW_STD = lambda g, v: 4 * g + v  # noqa: E731


def _to_layout(c, row=ROW_STD):
    C = np.zeros((16, 16))
    for lane in range(64):
        g, r = divmod(lane, 16)
        for v in range(4):
            C[row(g, v), r] = c[lane, v]
    return C


def _discover(c, want):
    """(g, v) -> the row of `want` that lane group g's element v holds."""
    found = {}
    for g in range(4):
        for v in range(4):
            got = np.array([c[16 * g + r, v] for r in range(16)])
            rows = [R for R in range(16) if np.allclose(got, want[R], rtol=1e-4, atol=1e-4)]
            found[(g, v)] = rows
    return found


def test_f64_16x16x4(probe):
    rng = np.random.default_rng(1)
    a, b = rng.standard_normal(64), rng.standard_normal(64)
    c = _run(probe, 0, a, b, np.zeros((64, 4)))
    want = np.einsum("gr,gc->rc", a.reshape(4, 16), b.reshape(4, 16))
    print("f64 rows", _discover(c, want))
    assert np.allclose(_to_layout(c, ROW_F64), want, rtol=1e-14, atol=1e-14)


def test_f32_16x16x4(probe):
    rng = np.random.default_rng(2)
    a = rng.standard_normal(64).astype(np.float32)
    b = rng.standard_normal(64).astype(np.float32)
    c = _run(probe, 1, a, b, np.zeros((64, 4), np.float32))
    want = np.einsum("gr,gc->rc", a.reshape(4, 16).astype(float), b.reshape(4, 16).astype(float))
    print("f32 rows", _discover(c, want))
    assert np.allclose(_to_layout(c), want, rtol=1e-5, atol=1e-6)


def _bf16_bits(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return r


def _bf16_val(bits):
    return (bits.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def test_bf16_16x16x32(probe):
    rng = np.random.default_rng(3)
    a = _bf16_bits(rng.standard_normal((64, 8)))
    b = _bf16_bits(rng.standard_normal((64, 8)))
    c = _run(probe, 2, a, b, np.zeros((64, 4), np.float32))
    av = _bf16_val(a).reshape(4, 16, 8)
    bv = _bf16_val(b).reshape(4, 16, 8)
    want = np.einsum("gre,gce->rc", av, bv)
    print("bf16 rows", _discover(c, want))
    assert np.allclose(_to_layout(c), want, rtol=1e-5, atol=1e-5)


def test_bf16_32x32x16(probe):
    """v_mfma_f32_32x32x16_bf16 (the adjoint's Gamma GEMM): lane l = 32 h + c
    supplies A[row c][k = 8h .. 8h + 7] and B[k = 8h ..][col c]; result element
    j of lane l is C[8 (j / 4) + 4 h + j % 4][c]."""
    rng = np.random.default_rng(5)
    a = _bf16_bits(rng.standard_normal((64, 8)))
    b = _bf16_bits(rng.standard_normal((64, 8)))
    c = _run(probe, 5, a, b, np.zeros((64, 16), np.float32))
    av = _bf16_val(a).reshape(2, 32, 8)
    bv = _bf16_val(b).reshape(2, 32, 8)
    want = np.einsum("hre,hce->rc", av, bv)
    got = np.zeros((32, 32))
    for lane in range(64):
        h, col = divmod(lane, 32)
        for j in range(16):
            got[8 * (j // 4) + 4 * h + j % 4, col] = c[lane, j]
    assert np.allclose(got, want, rtol=1e-5, atol=1e-5)


def e4m3_decode(codes):
    c = codes.astype(np.int64)
    s = np.where(c & 0x80, -1.0, 1.0)
    e = (c >> 3) & 0xF
    m = c & 7
    val = np.where(e == 0, m * 2.0 ** -9, (8 + m) * np.exp2(e - 10.0))
    return s * val


def _f8_k(g, e):
    """Hardware k of byte e of lane group g (v_mfma_scale_f32_16x16x128_f8f6f4)."""
    return 16 * g + e if e < 16 else 64 + 16 * g + (e - 16)


def _f8_expected(a, b, sa, sb):
    av = e4m3_decode(a).reshape(4, 16, 32)
    bv = e4m3_decode(b).reshape(4, 16, 32)
    want = np.zeros((16, 16))
    for g in range(4):
        for e in range(32):
            kb = _f8_k(g, e) // 32
            fa = np.exp2(sa[16 * kb + np.arange(16)] - 127.0)
            fb = np.exp2(sb[16 * kb + np.arange(16)] - 127.0)
            want += np.outer(av[g, :, e] * fa, bv[g, :, e] * fb)
    return want


@pytest.mark.parametrize("codeset", ["normal", "all"])
def test_f8_scaled_16x16x128(probe, codeset):
    """Format 0 is OCP E4M3; bytes 0-15 of lane group g are k = 16g.., bytes
    16-31 are k = 64 + 16g..; the scale of k block kb of row r is lane 16kb + r's."""
    rng = np.random.default_rng(4)
    codes = np.arange(256, dtype=np.uint8)
    if codeset == "normal":
        e = (codes >> 3) & 0xF
        valid = codes[(e >= 1) & (e <= 14)]
    else:
        valid = codes[(codes & 0x7F) != 0x7F]
    a = rng.choice(valid, size=(64, 32)).astype(np.uint8)
    b = rng.choice(valid, size=(64, 32)).astype(np.uint8)
    one = np.full(64, 127, np.int32)
    c = _run(probe, 3, a.view(np.int32), b.view(np.int32), np.zeros((64, 4), np.float32), one,
             one)
    want = _f8_expected(a, b, one, one)
    got = _to_layout(c)
    print("f8 uniform: got", got[0, :4], "want", want[0, :4], "nan", np.isnan(c).sum())
    print("f8 rows", _discover(c, want))
    assert np.allclose(got, want, rtol=1e-5, atol=1e-3 * np.abs(want).max())
    sa = rng.integers(120, 135, 64).astype(np.int32)
    sb = rng.integers(120, 135, 64).astype(np.int32)
    c = _run(probe, 3, a.view(np.int32), b.view(np.int32), np.zeros((64, 4), np.float32), sa, sb)
    want = _f8_expected(a, b, sa, sb)
    got = _to_layout(c)
    print("f8 scaled: got", got[0, :4], "want", want[0, :4])
    assert np.allclose(got, want, rtol=1e-5, atol=1e-3 * np.abs(want).max())


def test_i8_16x16x64(probe):
    """Byte e of lane 16 g + r of A meets byte e of lane 16 g + c of B (the
    kernels feed both from the same Q positions), and result element v of lane
    16 g + r is C[4 g + v][r], exactly in int32."""
    rng = np.random.default_rng(5)
    a = rng.integers(-128, 128, size=(64, 16)).astype(np.int8)
    b = rng.integers(-128, 128, size=(64, 16)).astype(np.int8)
    c = _run(probe, 4, a.view(np.int32), b.view(np.int32), np.zeros((64, 4), np.int32))
    av = a.astype(np.int64).reshape(4, 16, 16)
    bv = b.astype(np.int64).reshape(4, 16, 16)
    want = np.einsum("gre,gce->rc", av, bv)
    got = _to_layout(c.astype(np.float64))
    print("i8 rows", _discover(c.astype(np.float64), want.astype(np.float64)))
    assert np.array_equal(got, want.astype(np.float64))
