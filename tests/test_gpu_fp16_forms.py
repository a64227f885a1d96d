"""The packed FP16 projection's 16 operand forms (wl_store.h pk_cadd: one v_pk_add_f16 with
operand selections and negations per complex number), bit for bit against IEEE half-precision
addition in numpy, through the C ABI diagnostic wl_fp16_forms.  The projection they build is
HalfWilsonVector's (1 +- gamma_mu) psi (wilson_vector.h:439-567) in FP16 storage."""
import ctypes as C

import numpy as np
import pytest

from hila_amd import wilson as W

def pack(re, im):
    return (re.view(np.uint16).astype(np.uint32) | (im.view(np.uint16).astype(np.uint32) << 16))


def unpack(u):
    return (u & 0xFFFF).astype(np.uint16).view(np.float16), (u >> 16).astype(np.uint16).view(np.float16)


def reference(ar, ai, br, bi):
    """The 16 forms: for C = +1, -1, +i, -i the complex a + C b, returned as (re, im),
    (im, re), (re, -im), (im, -re); each half is one FP16 addition of the (negated) operands."""
    with np.errstate(over="ignore"):
        z = {  # C: (re terms, im terms) as pairs of FP16 addends
            0: ((ar, br), (ai, bi)),
            1: ((ar, -br), (ai, -bi)),
            2: ((ar, -bi), (ai, br)),
            3: ((ar, bi), (ai, -br)),
        }
        out = []
        for k in range(4):
            (r0, r1), (i0, i1) = z[k]
            re, im = r0 + r1, i0 + i1
            nre, nim = (-r0) + (-r1), (-i0) + (-i1)
            out += [(re, im), (im, re), (re, nim), (im, nre)]
    return out


def test_reference_known_values():
    """Pins the reference: 1 + 2i and 3 + 4i by hand, and the sign of a zero sum under -(x + y)."""
    h = lambda *v: np.array(v, dtype=np.float16)
    ref = reference(h(1), h(2), h(3), h(4))
    assert (ref[0][0][0], ref[0][1][0]) == (4, 6)          # a + b
    assert (ref[4][0][0], ref[4][1][0]) == (-2, -2)        # a - b
    assert (ref[8][0][0], ref[8][1][0]) == (-3, 5)         # a + i b = (1 - 4) + (2 + 3) i
    assert (ref[12][0][0], ref[12][1][0]) == (5, -1)       # a - i b
    assert (ref[11][0][0], ref[11][1][0]) == (5, 3)        # (im, -re) of a + i b
    z = reference(h(1), h(1), h(-1), h(-1))[2]             # (re, -im) of 1 + 1: im = 1 - 1
    assert z[1][0] == 0 and not np.signbit(z[1][0])        # (-1) + (+1) = +0


@pytest.mark.gpu
def test_fp16_operand_forms_bit_exact():
    rng = np.random.default_rng(11)
    n = 1 << 16
    bits = rng.integers(0, 1 << 16, size=(4, n), dtype=np.uint32).astype(np.uint16)
    vals = bits.view(np.float16)
    vals[~np.isfinite(vals)] = np.float16(0.5)  # finite inputs: NaN payloads are not compared
    # special values among the inputs: signed zeros, subnormals, the largest finite, ties
    specials = np.array([0.0, -0.0, 6e-8, -6e-8, 65504.0, -65504.0, 1.0, 2049.0, 0.5, -1.0], dtype=np.float16)
    vals[:, : specials.size] = specials
    vals[:, specials.size: 2 * specials.size] = specials[::-1]
    ar, ai, br, bi = vals
    a, b = pack(ar, ai), pack(br, bi)
    out = np.zeros(16 * n, dtype=np.uint32)
    rc = W.lib().wl_fp16_forms(0, a.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p), n,
                               out.ctypes.data_as(C.c_void_p))
    W.check(rc)
    ref = reference(ar, ai, br, bi)
    for f in range(16):
        got = out[f * n: (f + 1) * n]
        want = pack(np.asarray(ref[f][0], dtype=np.float16), np.asarray(ref[f][1], dtype=np.float16))
        bad = np.nonzero(got != want)[0]
        assert bad.size == 0, (f, bad[:5], [unpack(got[bad[:3]]), unpack(want[bad[:3]])])
