"""Double-double (about 106-bit) arithmetic in numpy, for checking residuals
below the FP64 floor on the CPU.

A double-double number is a pair (hi, lo) of float64 arrays with
|lo| <= ulp(hi)/2 and value hi + lo.  The error-free transforms are Knuth's
TwoSum and Dekker's TwoProduct with Veltkamp's split (Dekker, "A
floating-point technique for extending the available precision",
Numer. Math. 18, 1971), so no fused multiply-add is needed; numpy's
elementwise float64 operations round to nearest as IEEE 754 requires.

Complex double-double values are kept as four real arrays (re_hi, re_lo,
im_hi, im_lo) inside the small class CDD.  `mmdag_residual` evaluates
r = b - M M^dag (x_hi + x_lo) for the HubbardExp fermion matrix, given the
exact FP64 numbers of E and the phase table that the engine used, in
double-double throughout; `exact_mmdag_residual` does the same with Python
fractions (exact rational arithmetic) and pins this module on tiny inputs.

Test infrastructure only: nothing under nsl_amd/ imports it.
"""

from fractions import Fraction

import numpy as np

_SPLIT = 134217729.0  # 2^27 + 1


def two_sum(a, b):
    s = a + b
    bb = s - a
    err = (a - (s - bb)) + (b - bb)
    return s, err


def quick_two_sum(a, b):
    s = a + b
    return s, b - (s - a)


def split(a):
    t = _SPLIT * a
    hi = t - (t - a)
    return hi, a - hi


def two_prod(a, b):
    p = a * b
    ah, al = split(a)
    bh, bl = split(b)
    err = ((ah * bh - p) + ah * bl + al * bh) + al * bl
    return p, err


def dd_add(ah, al, bh, bl):
    s, e = two_sum(ah, bh)
    t, f = two_sum(al, bl)
    e = e + t
    s, e = quick_two_sum(s, e)
    e = e + f
    return quick_two_sum(s, e)


def dd_mul_d(ah, al, b):
    """(ah + al) * b for a float64 b."""
    p, e = two_prod(ah, b)
    e = e + al * b
    return quick_two_sum(p, e)


class CDD:
    """Complex double-double array."""

    def __init__(self, rh, rl, ih, il):
        self.rh, self.rl, self.ih, self.il = rh, rl, ih, il

    @classmethod
    def from_c128(cls, hi, lo=None):
        hi = np.asarray(hi, dtype=np.complex128)
        lo = np.zeros_like(hi) if lo is None else np.asarray(lo, dtype=np.complex128)
        rh, rl = quick_two_sum(hi.real.copy(), lo.real.copy())
        ih, il = quick_two_sum(hi.imag.copy(), lo.imag.copy())
        return cls(rh, rl, ih, il)

    def __add__(self, o):
        rh, rl = dd_add(self.rh, self.rl, o.rh, o.rl)
        ih, il = dd_add(self.ih, self.il, o.ih, o.il)
        return CDD(rh, rl, ih, il)

    def __neg__(self):
        return CDD(-self.rh, -self.rl, -self.ih, -self.il)

    def __sub__(self, o):
        return self + (-o)

    def mul_c128(self, z):
        """self * z for a complex128 array z (broadcast)."""
        zr = np.broadcast_to(z.real, self.rh.shape)
        zi = np.broadcast_to(z.imag, self.rh.shape)
        a = CDD(*dd_mul_d(self.rh, self.rl, zr), *dd_mul_d(self.ih, self.il, zr))     # (vr + i vi) zr
        b = CDD(*dd_mul_d(self.ih, self.il, -zi), *dd_mul_d(self.rh, self.rl, zi))    # (vr + i vi) i zi
        return a + b

    def scale(self, s):
        """self * s for a float64 array s broadcast against the value."""
        s = np.broadcast_to(s, self.rh.shape)
        return CDD(*dd_mul_d(self.rh, self.rl, s), *dd_mul_d(self.ih, self.il, s))

    def to_c128(self):
        return (self.rh + self.rl) + 1j * (self.ih + self.il)

    def norm2(self, axes):
        """sum |v|^2 over axes, as float64 (the value itself is tiny-error)."""
        v = self.to_c128()
        return (np.abs(v) ** 2).sum(axis=axes)


def _dd_rowsum(ph, pl):
    """Sum over the last axis in double-double: (sum_hi, sum_lo)."""
    sh = np.zeros(ph.shape[:-1])
    sl = np.zeros(ph.shape[:-1])
    for k in range(ph.shape[-1]):
        sh, sl = dd_add(sh, sl, ph[..., k], pl[..., k])
    return sh, sl


def _dd_matvec_real(E, vh, vl):
    """E @ v along the last axis of v (real double-double v), in
    double-double: out[..., y] = sum_x E[y, x] v[..., x]."""
    p, e = two_prod(E[None, :, :], vh.reshape(-1, 1, vh.shape[-1]))   # (rows, y, x)
    e = e + E[None, :, :] * vl.reshape(-1, 1, vl.shape[-1])
    sh, sl = _dd_rowsum(p, e * 0.0)
    # the low parts of the products and of v are summed in float64: they are
    # 2^-53 below the products, so this adds an error of about 2^-106 * n
    sl = sl + e.sum(axis=-1)
    sh, sl = quick_two_sum(sh, sl)
    return sh.reshape(vh.shape), sl.reshape(vh.shape)


def dd_E_apply(E, v):
    """Complex double-double v (..., nx): out = v @ E^T (E real)."""
    rh, rl = _dd_matvec_real(E, v.rh, v.rl)
    ih, il = _dd_matvec_real(E, v.ih, v.il)
    return CDD(rh, rl, ih, il)


def _roll(v, shift):
    return CDD(*(np.roll(a, shift, axis=-2) for a in (v.rh, v.rl, v.ih, v.il)))


def dd_M(E, P, v, field_of):
    """(M v)_t = v_t - b_t E (P_{t-1} v_{t-1}), b_0 = -1; v (nb, nt, nx)."""
    nt = v.rh.shape[-2]
    Pb = P[field_of]                                   # (nb, nt, nx)
    sh = _roll(v.mul_c128(Pb), 1)                      # P_{t-1} v_{t-1}
    ev = dd_E_apply(E, sh)
    sign = np.ones((1, nt, 1))
    sign[0, 0, 0] = -1.0
    return v - ev.scale(sign)


def dd_Mdag(E, P, v, field_of):
    """(M^dag v)_t = v_t - conj(P_t) b'_t E v_{t+1}, b'_{nt-1} = -1."""
    nt = v.rh.shape[-2]
    ev = dd_E_apply(E, _roll(v, -1))
    sign = np.ones((1, nt, 1))
    sign[0, nt - 1, 0] = -1.0
    return v - ev.mul_c128(np.conj(P[field_of])).scale(sign)


def mmdag_residual(E, P, b, x_hi, x_lo, rpf=1, op="MMdagger"):
    """r = b - M M^dag (x_hi + x_lo) (op "MMdagger") or b - M^dag M (x_hi +
    x_lo) (op "MdaggerM") in double-double; returns (r as complex128, per-RHS
    relative residual |r| / |b|).  E: (nx, nx) float64; P: (nfield, nt, nx)
    complex128; b, x_hi, x_lo: (nb, nt, nx) complex128."""
    nb = b.shape[0]
    field_of = np.arange(nb) // rpf
    x = CDD.from_c128(x_hi, x_lo)
    if op == "MMdagger":
        t = dd_M(E, P, dd_Mdag(E, P, x, field_of), field_of)
    else:
        t = dd_Mdag(E, P, dd_M(E, P, x, field_of), field_of)
    r = CDD.from_c128(b) - t
    rc = r.to_c128()
    rel = np.sqrt((np.abs(rc) ** 2).sum(axis=(1, 2)) / (np.abs(b) ** 2).sum(axis=(1, 2)))
    return rc, rel


# ---------------------------------------------------------------- exact check
def _fc(z):
    return (Fraction(float(z.real)), Fraction(float(z.imag)))


def exact_mmdag_residual(E, P, b, x_hi, x_lo, op="MMdagger"):
    """The same residual for ONE right-hand side (b, x: (nt, nx)) and one
    field (P: (nt, nx)) in exact rational arithmetic; returns complex128 r
    (each component rounded once).  Slow: tiny lattices only."""
    nt, nx = b.shape
    Ef = [[Fraction(float(E[y, x])) for x in range(nx)] for y in range(nx)]
    Pf = [[_fc(P[t, x]) for x in range(nx)] for t in range(nt)]
    X = [[(Fraction(float(x_hi[t, i].real)) + Fraction(float(x_lo[t, i].real)),
           Fraction(float(x_hi[t, i].imag)) + Fraction(float(x_lo[t, i].imag))) for i in range(nx)]
         for t in range(nt)]

    def cmul(a, c):
        return (a[0] * c[0] - a[1] * c[1], a[0] * c[1] + a[1] * c[0])

    def eapply(v):
        return [(sum(Ef[y][x] * v[x][0] for x in range(nx)), sum(Ef[y][x] * v[x][1] for x in range(nx)))
                for y in range(nx)]

    def mdag(v):   # (M^dag v)_t = v_t - b'_t conj(P_t) E v_{t+1}
        out = []
        for t in range(nt):
            ev = eapply(v[(t + 1) % nt])
            sgn = -1 if t == nt - 1 else 1
            row = []
            for i in range(nx):
                q = cmul((Pf[t][i][0], -Pf[t][i][1]), ev[i])
                row.append((v[t][i][0] - sgn * q[0], v[t][i][1] - sgn * q[1]))
            out.append(row)
        return out

    def m(v):      # (M v)_t = v_t - b_t E (P_{t-1} v_{t-1})
        out = []
        for t in range(nt):
            ev = eapply([cmul(Pf[(t - 1) % nt][i], v[(t - 1) % nt][i]) for i in range(nx)])
            sgn = -1 if t == 0 else 1
            out.append([(v[t][i][0] - sgn * ev[i][0], v[t][i][1] - sgn * ev[i][1]) for i in range(nx)])
        return out

    tv = m(mdag(X)) if op == "MMdagger" else mdag(m(X))
    r = np.zeros((nt, nx), dtype=np.complex128)
    for t in range(nt):
        for i in range(nx):
            br, bi = _fc(b[t, i])
            r[t, i] = complex(float(br - tv[t][i][0]), float(bi - tv[t][i][1]))
    return r
