"""The oracle's 16-bit roundings against torch's own casts (an independent
implementation: c10's BFloat16 / Half conversions).

``oracle.spmv.round_bf16`` and ``oracle.fsai._bf16`` are what every BF16 parity
check of the library rests on (the 16-bit SpMV copy, BF16 FSAI factors, the
edge-MLP's matrix-core operands), so they are compared bit for bit with
``torch.Tensor.to(torch.bfloat16)`` on random values of every magnitude, exact
ties with an odd and an even kept bit, +-0, +-Inf, NaN payloads, subnormals and
values next to FLT_MAX; the FP16 rounding used by ``round16_rowsum`` is
compared with ``torch.float16``. Planted mistakes (truncation, ties away from
zero, the unguarded NaN add) fail the same comparisons.
"""
import numpy as np
import pytest
import torch

from oracle import fsai as OF
from oracle import spmv as OS


def _bits(u):
    return np.asarray(u, dtype=np.uint32).view(np.float32)


def _cases():
    rng = np.random.default_rng(11)
    parts = [
        # random bit patterns cover every exponent, both signs, NaNs and subnormals
        rng.integers(0, 2 ** 32, size=200_000, dtype=np.uint64).astype(np.uint32),
        # ordinary magnitudes
        rng.standard_normal(50_000).astype(np.float32).view(np.uint32),
        (rng.standard_normal(50_000) * 1e-30).astype(np.float32).view(np.uint32),
        (rng.standard_normal(50_000) * 1e30).astype(np.float32).view(np.uint32),
    ]
    hi = rng.integers(0, 2 ** 16, size=20_000, dtype=np.uint64).astype(np.uint32) << 16
    parts += [
        hi | 0x8000,            # exact ties: kept bit (bit 16) odd or even
        hi | 0x7FFF, hi | 0x8001,  # just below / above a tie
        hi | 0xFFFF, hi | 0x0001, hi,
    ]
    special = [0x00000000, 0x80000000,  # +-0
               0x7F800000, 0xFF800000,  # +-Inf
               0x7FC00000, 0xFFC00000, 0x7F800001, 0xFF800001, 0x7FFFFFFF, 0xFFFFFFFF,
               0x7FBFFFFF, 0x7F808000, 0x7FFF8000,  # NaN payloads (signalling and quiet)
               0x00000001, 0x80000001, 0x00008000, 0x00018000, 0x007FFFFF, 0x807FFFFF,
               0x00800000, 0x007F8000, 0x007F7FFF,  # subnormals and the smallest normal
               0x7F7FFFFF, 0xFF7FFFFF, 0x7F7F8000, 0x7F7F7FFF, 0x7F7F0000, 0x7F7E8000,  # near FLT_MAX
               0x3F808000, 0x3F818000, 0x3F807FFF, 0x3F80FFFF,  # ties around 1.0
               ]
    parts.append(np.array(special, dtype=np.uint32))
    return _bits(np.concatenate(parts))


def _torch_round(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dtype).float().numpy()


def _same(got, want):
    """Bit-equal, except that NaN must match NaN (torch canonicalises the payload)."""
    gn, wn = np.isnan(got), np.isnan(want)
    assert np.array_equal(gn, wn), np.flatnonzero(gn != wn)[:10]
    gb, wb = got.view(np.uint32)[~gn], want.view(np.uint32)[~wn]
    bad = np.flatnonzero(gb != wb)
    assert bad.size == 0, [(hex(int(a)), hex(int(b))) for a, b in zip(gb[bad[:5]], wb[bad[:5]])]


def test_round_bf16_matches_torch():
    a = _cases()
    _same(OS.round_bf16(a), _torch_round(a, torch.bfloat16))


def test_round_bf16_nan_stays_nan_and_keeps_sign():
    a = _bits([0x7FFFFFFF, 0xFFFFFFFF, 0x7F800001, 0xFF800001])
    r = OS.round_bf16(a)
    assert np.isnan(r).all()
    assert list(np.signbit(r)) == [False, True, False, True]
    # the device converter keeps (bits >> 16) | 0x40: a quiet NaN with the high payload
    assert [hex(int(u) >> 16) for u in r.view(np.uint32)] == ["0x7fff", "0xffff", "0x7fc0", "0xffc0"]


def test_fsai_bf16_matches_torch():
    a = _cases()
    got = OF._bf16(a, True)
    assert got.dtype == np.float64
    _same(got.astype(np.float32), _torch_round(a, torch.bfloat16))
    # from float64 input: one FP32 rounding first, as the kernels store FP32 operands
    d = np.random.default_rng(3).standard_normal(10_000) * 10.0 ** np.random.default_rng(4).integers(-20, 20, 10_000)
    _same(OF._bf16(d, True).astype(np.float32), _torch_round(d.astype(np.float32), torch.bfloat16))
    assert np.array_equal(OF._bf16(d, False), d)


def test_fp16_rounding_matches_torch():
    """round16_rowsum's FP16 rounding (numpy float16) against torch.float16."""
    a = _cases()
    with np.errstate(over="ignore"):  # overflow to +-Inf is one of the cases
        got = a.astype(np.float16).astype(np.float32)
    _same(got, _torch_round(a, torch.float16))


def test_round16_rowsum_offdiagonals_match_torch():
    """The 16-bit value copy's off-diagonals are torch's roundings of the FP32 values."""
    rng = np.random.default_rng(5)
    n = 300
    rows = np.repeat(np.arange(n), 5)
    cols = np.clip(rows + np.tile(np.arange(-2, 3), n), 0, n - 1)
    order = np.lexsort((cols, rows))
    rows, cols = rows[order], cols[order]
    vals = (rng.standard_normal(rows.size) * 3).astype(np.float32)
    rp = np.searchsorted(rows, np.arange(n + 1)).astype(np.int64)
    for fmt, dt in (("bf16", torch.bfloat16), ("f16", torch.float16)):
        v16, _ = OS.round16_rowsum(rp, cols, vals, fmt=fmt)
        off = cols != rows
        _same(v16[off].astype(np.float32), _torch_round(vals[off], dt))


@pytest.mark.parametrize("mistake", ["truncate", "ties_away", "unguarded_nan"])
def test_planted_mistakes_are_caught(mistake):
    """Plausible mistakes in a BF16 rounding differ from torch on these cases."""
    a = _cases()
    u = a.view(np.uint32).astype(np.uint64)
    if mistake == "truncate":
        r = (u >> 16).astype(np.uint32)
    elif mistake == "ties_away":
        r = ((u + 0x8000) >> 16).astype(np.uint32)
    else:
        r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32)
    bad = (r.astype(np.uint32) << 16).view(np.float32)
    with pytest.raises(AssertionError):
        _same(bad, _torch_round(a, torch.bfloat16))
