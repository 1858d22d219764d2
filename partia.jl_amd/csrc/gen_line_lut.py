#!/usr/bin/env python3
"""Generate the line-integrated SPH kernel tables used by the HIP library.

For each 3D kernel shape f(q) with support radius K, the table holds

    I(q_perp) = 2 * integral_0^{sqrt(K^2 - q_perp^2)} f(sqrt(q_perp^2 + s^2)) ds

on the 4096-point grid Q_k = k*K/4095 (k = 0..4095, evaluated as (k*K)/4095 in
float64, the same grid as src/julia/KernelInterpolation/table/
line_integrated_kernel_tables.jl:1), and the normalised table Inorm = I * sigma_3
(sigma_3 = the 3D normalisation of kernel_function/kernels/*.jl).

The integral is evaluated here from the kernel definitions, not copied: the
s-interval is split at every breakpoint of the piecewise polynomial f (where
sqrt(q_perp^2 + s^2) crosses a knot), the segment touching s = 0 is graded
geometrically towards 0 (the integrand's branch points sit at s = +-i q_perp),
and each piece is integrated with 48-point Gauss-Legendre in 80-bit long double.
The result is rounded once to float64.

Usage: gen_line_lut.py OUT.inc   (writes a C/HIP include with the tables)
"""
import sys
import numpy as np

LD = np.longdouble
NPOINT = 4096
NGL = 48

# (name, support K, knots strictly inside (0, K), sigma_3 numerator/denominator)
# sigma_3 is evaluated exactly as the reference writes it in float64:
#   M4 1/pi, M5 0.05/pi, M6 1/(120 pi), C2 21/(16 pi), C4 495/(256 pi), C6 1365/(512 pi)
KERNELS = [
    ("M4_spline", 2.0, (1.0,)),
    ("M5_spline", 2.5, (0.5, 1.5)),
    ("M6_spline", 3.0, (1.0, 2.0)),
    ("C2_Wendland", 2.0, ()),
    ("C4_Wendland", 2.0, ()),
    ("C6_Wendland", 2.0, ()),
]


def sigma3(name):
    pi = np.float64(np.pi)
    return {
        "M4_spline": np.float64(1.0) / pi,
        "M5_spline": np.float64(0.05) / pi,
        "M6_spline": np.float64(1.0) / (np.float64(120.0) * pi),
        "C2_Wendland": np.float64(21.0) / (np.float64(16.0) * pi),
        "C4_Wendland": np.float64(495.0) / (np.float64(256.0) * pi),
        "C6_Wendland": np.float64(1365.0) / (np.float64(512.0) * pi),
    }[name]


def shape(name, q):
    """Dimensionless kernel shape f(q) (q >= 0) in long double, piecewise as in
    kernel_function/kernels/<name>.jl."""
    q = np.asarray(q, dtype=LD)
    one, two = LD(1), LD(2)
    out = np.zeros_like(q)
    if name == "M4_spline":
        a = np.where(q < 2, (two - q) ** 3 * LD(0.25), 0)
        b = np.where(q < 1, (one - q) ** 3, 0)
        out = a - b
    elif name == "M5_spline":
        a = np.where(q < 2.5, (LD(2.5) - q) ** 4, 0)
        b = np.where(q < 1.5, LD(5) * (LD(1.5) - q) ** 4, 0)
        c = np.where(q < 0.5, LD(10) * (LD(0.5) - q) ** 4, 0)
        out = a - b + c
    elif name == "M6_spline":
        a = np.where(q < 3, (LD(3) - q) ** 5, 0)
        b = np.where(q < 2, LD(6) * (two - q) ** 5, 0)
        c = np.where(q < 1, LD(15) * (one - q) ** 5, 0)
        out = a - b + c
    elif name == "C2_Wendland":
        out = np.where(q < 2, (one - q / 2) ** 4 * (2 * q + 1), 0)
    elif name == "C4_Wendland":
        out = np.where(q < 2, (one - q / 2) ** 6 * (LD(35) / LD(12) * q * q + 3 * q + 1), 0)
    elif name == "C6_Wendland":
        out = np.where(q < 2, (one - q / 2) ** 8 * (4 * q ** 3 + LD(6.25) * q * q + 4 * q + 1), 0)
    return out


def gauss_legendre_ld(n):
    """n-point Gauss-Legendre nodes/weights on [-1, 1], Newton-refined in long double."""
    x64, _ = np.polynomial.legendre.leggauss(n)
    x = x64.astype(LD)
    for _ in range(6):
        p0, p1 = np.ones_like(x), x.copy()
        for k in range(2, n + 1):
            p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
        dp = n * (x * p1 - p0) / (x * x - 1)
        x = x - p1 / dp
    p0, p1 = np.ones_like(x), x.copy()
    for k in range(2, n + 1):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
    dp = n * (x * p1 - p0) / (x * x - 1)
    w = 2 / ((1 - x * x) * dp * dp)
    return x, w


GX, GW = gauss_legendre_ld(NGL)


def _segment(name, qp, a, b):
    if b <= a:
        return LD(0)
    mid, half = (a + b) / 2, (b - a) / 2
    s = mid + half * GX
    r = np.sqrt(qp * qp + s * s)
    return half * np.sum(GW * shape(name, r))


def line_integral(name, K, knots, qp):
    qp = LD(qp)
    KL = LD(K)
    if qp >= KL:
        return LD(0)
    smax = np.sqrt(KL * KL - qp * qp)
    cuts = [LD(0)]
    for kn in knots:
        kn = LD(kn)
        if kn > qp:
            cuts.append(np.sqrt(kn * kn - qp * qp))
    cuts.append(smax)
    total = LD(0)
    for i in range(len(cuts) - 1):
        a, b = cuts[i], cuts[i + 1]
        if i == 0 and qp > 0:
            # geometric grading towards s = 0, scale set by q_perp
            edges = [a]
            e = min(qp, b)
            while e < b:
                edges.append(e)
                e = e * 2
            edges.append(b)
            for j in range(len(edges) - 1):
                total += _segment(name, qp, edges[j], edges[j + 1])
        else:
            total += _segment(name, qp, a, b)
    return 2 * total


def tables(name, K, knots):
    k = np.arange(NPOINT, dtype=np.float64)
    Q = (k * np.float64(K)) / np.float64(NPOINT - 1)
    I = np.array([line_integral(name, K, knots, q) for q in Q], dtype=LD).astype(np.float64)
    I[-1] = 0.0
    Inorm = I * sigma3(name)
    return Q, I, Inorm


def main(out_path):
    lines = ["// Generated by gen_line_lut.py: line-integrated kernel tables (float64).",
             "// Order of kernels: M4, M5, M6, C2, C4, C6.  Q_k = (k*K)/4095.",
             "#define PH_LUT_N %d" % NPOINT,
             "static const double ph_lut_support[6] = {2.0, 2.5, 3.0, 2.0, 2.0, 2.0};",
             "static const double ph_lut_inorm[6][%d] = {" % NPOINT]
    for name, K, knots in KERNELS:
        _, _, Inorm = tables(name, K, knots)
        body = ",".join(float(v).hex() for v in Inorm)
        lines.append("{%s}," % body.replace("0x", "0x"))
    lines.append("};")
    text = "\n".join(lines) + "\n"
    # hex float literals are exact and valid in C99/C++17
    with open(out_path, "w") as f:
        f.write(text)


if __name__ == "__main__":
    main(sys.argv[1])
