"""Special-function tables of the nAiT density functional (build-time product).

Writes mpdpft_amd/lib/nait_tables.bin, read by the HIP library at context
creation.  Three tables, all computed here from the mathematics:

* CurlyA(y) = integral_y^inf Ai(t) dt at the knots y_i = -150 + 0.01 i
  (i = 0..25000) -- the knots of the reference's TabFunc_CurlyA.dat
  (mpDPFT.h:231-237 evaluates a cubic spline through them).  Values by
  composite 20-point Gauss-Legendre quadrature of Ai over each 0.01 panel,
  summed from the right (CurlyA(100) from [100, 115]).
* Ai(y), Ai'(y) on [-300, 100], step 0.005 (for the Airy energy integral);
  quintic Hermite interpolation with Ai'' = y Ai.
* L_s(x) = Li_s(-e^x) for s = 3/2, 1/2, -1/2, -3/2, -5/2, -7/2 on [-40, 200],
  step 0.01.  d/dx L_s = L_{s-1}, so each function's quintic Hermite
  interpolant uses the next two of the chain.  Values from
  L_{3/2-m}(x) = -(1/Gamma(3/2)) int_0^inf t^{1/2} sigma^{(m)}(x - t) dt,
  sigma(z) = 1/(1+e^-z), in extended precision (t = u^2, composite
  Gauss-Legendre).

Run:  python -m mpdpft_amd.tables   (or via mpdpft_amd/build.py)
"""
import hashlib
import math
import os
import struct
import sys

import numpy as np
from scipy import special as sp

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "lib", "nait_tables.bin")
MAGIC = b"NAITTAB1"

C_X0, C_H, C_N = -150.0, 0.01, 25001
A_X0, A_H, A_N = -300.0, 0.005, 80001
P_X0, P_H, P_N = -40.0, 0.01, 24001
P_ORDERS = (1.5, 0.5, -0.5, -1.5, -2.5, -3.5)


def _gl(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return x, w


def curly_a_knots():
    """CurlyA at the 25001 knots, by panel quadrature of Ai summed from the right."""
    xs = C_X0 + C_H * np.arange(C_N)
    gx, gw = _gl(20)
    a = xs[:-1]
    b = xs[1:]
    mid, half = 0.5 * (a + b), 0.5 * (b - a)
    t = mid[:, None] + half[:, None] * gx[None, :]
    ai = sp.airy(t)[0]
    panels = (ai * gw[None, :]).sum(1) * half
    # tail beyond the last knot
    tt = np.linspace(xs[-1], xs[-1] + 15.0, 1501)
    ta, tb = tt[:-1], tt[1:]
    tm, th = 0.5 * (ta + tb), 0.5 * (tb - ta)
    tail = (sp.airy(tm[:, None] + th[:, None] * gx[None, :])[0] * gw[None, :]).sum(1) * th
    c = np.empty(C_N)
    c[-1] = tail[::-1].sum()  # small terms first
    acc = c[-1]
    for i in range(C_N - 2, -1, -1):
        acc = acc + panels[i]
        c[i] = acc
    return xs, c


def airy_table():
    xs = A_X0 + A_H * np.arange(A_N)
    ai, aip, _, _ = sp.airy(xs)
    return xs, np.stack([ai, aip], 1)


def _sigma_derivs(z):
    """sigma^{(m)}(z), m = 0..5, from p = sigma(z), q = sigma(-z) (both accurate)."""
    e = np.exp(-np.abs(z))
    one = np.ones_like(z)
    pbig = one / (one + e)      # sigma(|z|)
    psmall = e / (one + e)      # sigma(-|z|)
    p = np.where(z >= 0, pbig, psmall)
    q = np.where(z >= 0, psmall, pbig)
    pq = p * q
    d = [p, pq, pq * (q - p), pq * (1 - 6 * pq), pq * (q - p) * (1 - 12 * pq),
         pq * (1 - 30 * pq + 120 * pq * pq)]
    return d


def polylog_chain_quadrature(x, nodes=24, du=0.05):
    """L_{3/2-m}(x) for m = 0..5 at points x (extended precision quadrature)."""
    x = np.asarray(x, dtype=np.longdouble)
    gx, gw = _gl(nodes)
    gx = gx.astype(np.longdouble)
    gw = gw.astype(np.longdouble)
    out = np.zeros((len(x), 6), dtype=np.longdouble)
    g32 = np.longdouble(math.gamma(1.5))
    for i, xi in enumerate(x):
        U = math.sqrt(max(float(xi), 0.0) + 60.0)
        npan = int(math.ceil(U / du))
        edges = np.linspace(0, U, npan + 1).astype(np.longdouble)
        a, b = edges[:-1], edges[1:]
        mid, half = (a + b) / 2, (b - a) / 2
        u = (mid[:, None] + half[:, None] * gx[None, :]).reshape(-1)
        w = (half[:, None] * gw[None, :]).reshape(-1)
        z = xi - u * u
        ds = _sigma_derivs(z)
        base = 2 * u * u * w
        for m in range(6):
            out[i, m] = -(base * ds[m]).sum() / g32
    return out


def eta(s):
    """Dirichlet eta via scipy's Riemann zeta."""
    if s == 1:
        return math.log(2.0)
    return (1 - 2.0 ** (1 - s)) * float(sp.zeta(s))


def polylog_taylor(s, x, terms=90):
    """Li_s(-e^x) = -sum_n eta(s-n) x^n / n!  (|x| < pi)."""
    x = np.asarray(x, dtype=float)
    acc = np.zeros_like(x)
    for n in range(terms):
        acc += eta(s - n) * x ** n / math.factorial(n)
    return -acc


def polylog_series(s, x, terms=60):
    """Li_s(-e^x) = sum_k (-1)^k e^{kx} / k^s  (x < 0)."""
    x = np.asarray(x, dtype=float)
    acc = np.zeros_like(x)
    for k in range(terms, 0, -1):
        acc += (-1.0) ** k * np.exp(k * x) / k ** s
    return acc


def polylog_asymptotic(s, x, K=12):
    """Sommerfeld expansion -x^s/Gamma(s+1) sum_k 2 eta(2k) (s)_{2k} x^{-2k}, x >> 1."""
    x = np.asarray(x, dtype=float)
    acc = np.ones_like(x)
    ff = 1.0
    for k in range(1, K + 1):
        ff *= (s - 2 * k + 2) * (s - 2 * k + 1)
        acc += 2 * eta(2 * k) * ff * x ** (-2.0 * k)
    return -x ** s / math.gamma(s + 1) * acc


def asymptotic_coefficients(K=12):
    """2 eta(2k) for k = 1..K (used by the runtime asymptotic branch)."""
    return [2 * eta(2 * k) for k in range(1, K + 1)]


def _source_hash():
    with open(os.path.abspath(__file__), "rb") as f:
        return hashlib.sha256(f.read()).digest()[:8]


def write(path=OUT, verbose=True):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    xc, c = curly_a_knots()
    xa, a = airy_table()
    xp = P_X0 + P_H * np.arange(P_N)
    P = polylog_chain_quadrature(xp).astype(np.float64)
    with open(path + ".tmp", "wb") as f:
        f.write(MAGIC)
        f.write(_source_hash())
        f.write(struct.pack("<qdd", C_N, C_X0, C_H))
        f.write(struct.pack("<qdd", A_N, A_X0, A_H))
        f.write(struct.pack("<qdd", P_N, P_X0, P_H))
        f.write(np.ascontiguousarray(c, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(P, dtype="<f8").tobytes())
    os.replace(path + ".tmp", path)
    if verbose:
        print("wrote", path)
    return path


def read(path=OUT):
    with open(path, "rb") as f:
        buf = f.read()
    assert buf[:8] == MAGIC
    off = 16
    cn, cx0, ch = struct.unpack_from("<qdd", buf, off)
    off += 24
    an, ax0, ah = struct.unpack_from("<qdd", buf, off)
    off += 24
    pn, px0, ph = struct.unpack_from("<qdd", buf, off)
    off += 24
    c = np.frombuffer(buf, "<f8", cn, off)
    off += 8 * cn
    a = np.frombuffer(buf, "<f8", 2 * an, off).reshape(an, 2)
    off += 16 * an
    p = np.frombuffer(buf, "<f8", 6 * pn, off).reshape(pn, 6)
    return dict(c=c, c_x0=cx0, c_h=ch, a=a, a_x0=ax0, a_h=ah, p=p, p_x0=px0, p_h=ph,
                hash=buf[8:16])


def up_to_date(path=OUT):
    if not os.path.exists(path):
        return False
    with open(path, "rb") as f:
        head = f.read(16)
    return head[:8] == MAGIC and head[8:16] == _source_hash()


def ensure(path=OUT, verbose=True):
    if not up_to_date(path):
        write(path, verbose)
    return path


if __name__ == "__main__":
    write()
    sys.exit(0)
