"""OCP 8-bit floating point E4M3 ("FN": no infinities, one NaN code per sign)
as defined by the OCP 8-bit Floating Point Specification (OFP8) v1.0: sign,
4 exponent bits with bias 7, 3 mantissa bits; exponent field 0 holds the
subnormals m/8 * 2^-6; the largest finite value is 448 (S.1111.110) and
S.1111.111 is NaN.  Encoding rounds to nearest, ties to even."""
import numpy as np


def _table():
    vals = np.empty(256)
    for code in range(256):
        s = -1.0 if code & 0x80 else 1.0
        e = (code >> 3) & 0xF
        m = code & 0x7
        if e == 0xF and m == 0x7:
            vals[code] = np.nan
        elif e == 0:
            vals[code] = s * m / 8.0 * 2.0 ** -6
        else:
            vals[code] = s * (1.0 + m / 8.0) * 2.0 ** (e - 7)
    return vals


DECODE = _table()
_POS = np.array([c for c in range(0x7F)])          # finite non-negative codes 0x00 .. 0x7E
_POSV = DECODE[_POS]


def decode(codes: np.ndarray) -> np.ndarray:
    return DECODE[np.asarray(codes, dtype=np.uint8)]


def encode(x: np.ndarray) -> np.ndarray:
    """Round-to-nearest-even encoding of finite |x| <= 448."""
    x = np.asarray(x, dtype=np.float64)
    a = np.abs(x)
    assert np.all(a <= 448.0)
    idx = np.searchsorted(_POSV, a)            # first grid value >= a
    idx = np.clip(idx, 1, len(_POSV) - 1)
    lo, hi = _POSV[idx - 1], _POSV[idx]
    pick_hi = (a - lo > hi - a) | ((a - lo == hi - a) & (_POS[idx] % 2 == 0))
    code = np.where(a == 0, 0, np.where(pick_hi, _POS[idx], _POS[idx - 1]))
    code = np.where(a >= 448.0, 0x7E, code)
    sign = np.signbit(x).astype(np.uint8) << 7
    return (code.astype(np.uint8) | sign).astype(np.uint8)


def encode_saturating(x: np.ndarray) -> np.ndarray:
    """encode() after saturating |x| > 448 (infinities included) to +-448, the conversion of the FP8 Conv1d
    kernels; NaN is left to the caller."""
    x = np.asarray(x, dtype=np.float64)
    return encode(np.where(np.abs(x) > 448.0, np.copysign(448.0, x), x))
