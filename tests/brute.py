"""O(N) brute-force references in numpy, independent of any tree.

Restates the reference's test references test/interpolation_test_common.jl:31-351
(brute_density, brute_quantity, brute_gradient_quantity, brute_divergence,
brute_curl, line-integrated density/quantity) and test/grid_interpolation_test_common.jl:242-260
(brute_nearest_h), with kernel shapes written from their closed forms (not from
the oracle's code).  Sums run in numpy's pairwise order, so agreement with the
tree-based paths is to rounding (rtol ~1e-10), as in the reference tests.
"""
import numpy as np

SUPPORT = {0: 2.0, 1: 2.5, 2: 3.0, 3: 2.0, 4: 2.0, 5: 2.0}
SIGMA3 = {0: 1 / np.pi, 1: 0.05 / np.pi, 2: 1 / (120 * np.pi), 3: 21 / (16 * np.pi), 4: 495 / (256 * np.pi),
          5: 1365 / (512 * np.pi)}


def shape(K, q):
    q = np.asarray(q, dtype=float)
    if K == 0:
        return np.where(q < 1, 0.25 * (2 - q) ** 3 - (1 - q) ** 3, np.where(q < 2, 0.25 * (2 - q) ** 3, 0.0))
    if K == 1:
        a = np.where(q < 2.5, (2.5 - q) ** 4, 0.0)
        b = np.where(q < 1.5, 5 * (1.5 - q) ** 4, 0.0)
        c = np.where(q < 0.5, 10 * (0.5 - q) ** 4, 0.0)
        return a - b + c
    if K == 2:
        a = np.where(q < 3, (3 - q) ** 5, 0.0)
        b = np.where(q < 2, 6 * (2 - q) ** 5, 0.0)
        c = np.where(q < 1, 15 * (1 - q) ** 5, 0.0)
        return a - b + c
    w = np.clip(1 - 0.5 * q, 0, None)
    if K == 3:
        return np.where(q < 2, w ** 4 * (2 * q + 1), 0.0)
    if K == 4:
        return np.where(q < 2, w ** 6 * (35 / 12 * q * q + 3 * q + 1), 0.0)
    return np.where(q < 2, w ** 8 * (4 * q ** 3 + 6.25 * q * q + 4 * q + 1), 0.0)


def dshape(K, q, eps=None):
    """f'(q), closed form."""
    q = np.asarray(q, dtype=float)
    if K == 0:
        return np.where(q < 1, -0.75 * (2 - q) ** 2 + 3 * (1 - q) ** 2, np.where(q < 2, -0.75 * (2 - q) ** 2, 0.0))
    if K == 1:
        a = np.where(q < 2.5, -4 * (2.5 - q) ** 3, 0.0)
        b = np.where(q < 1.5, 20 * (1.5 - q) ** 3, 0.0)
        c = np.where(q < 0.5, -40 * (0.5 - q) ** 3, 0.0)
        return a + b + c
    if K == 2:
        a = np.where(q < 3, -5 * (3 - q) ** 4, 0.0)
        b = np.where(q < 2, 30 * (2 - q) ** 4, 0.0)
        c = np.where(q < 1, -75 * (1 - q) ** 4, 0.0)
        return a + b + c
    w = np.clip(1 - 0.5 * q, 0, None)
    if K == 3:
        return np.where(q < 2, -5 * q * w ** 3, 0.0)  # d/dq (1-q/2)^4 (2q+1) = -5 q (1-q/2)^3
    if K == 4:
        return np.where(q < 2, w ** 6 * (35 / 6 * q + 3) - 3 * w ** 5 * (35 / 12 * q * q + 3 * q + 1), 0.0)
    return np.where(q < 2, w ** 8 * (12 * q * q + 12.5 * q + 4) - 4 * w ** 7 * (4 * q ** 3 + 6.25 * q * q + 4 * q + 1),
                    0.0)


def W(K, r, h):
    return SIGMA3[K] * shape(K, r / h) / h ** 3


def gradW(K, dx, dy, dz, h):
    r = np.sqrt(dx * dx + dy * dy + dz * dz)
    with np.errstate(invalid="ignore", divide="ignore"):
        c = np.where(r > 0, SIGMA3[K] * dshape(K, r / h) / (h ** 4 * np.where(r > 0, r, 1.0)), 0.0)
    return dx * c, dy * c, dz * c


def _weights(inp, p, ha, K, strategy, with_grad):
    x, y, z = inp["x"], inp["y"], inp["z"]
    dx, dy, dz = p[0] - x, p[1] - y, p[2] - z
    d2 = dx * dx + dy * dy + dz * dz
    hb = inp["h"]
    Kv = SUPPORT[K]
    if strategy == 0:
        rad = Kv * ha * np.ones_like(hb)
    elif strategy == 1:
        rad = Kv * hb
    else:
        rad = Kv * np.maximum(ha, hb)
    sel = d2 <= rad * rad
    r = np.sqrt(d2[sel])
    hsel = hb[sel]
    if strategy == 2:
        w = 0.5 * (W(K, r, ha) + W(K, r, hsel))
    else:
        w = W(K, r, ha if strategy == 0 else hsel)
    g = None
    if with_grad:
        a = (dx[sel], dy[sel], dz[sel])
        if strategy == 2:
            ga = gradW(K, *a, ha)
            gb = gradW(K, *a, hsel)
            g = tuple(0.5 * (u + v) for u, v in zip(ga, gb))
        else:
            g = gradW(K, *a, ha if strategy == 0 else hsel)
    return sel, w, g


def density(inp, p, ha, K, strategy):
    sel, w, _ = _weights(inp, p, ha, K, strategy, False)
    return float(np.sum(inp["m"][sel] * w))


def quantity(inp, p, ha, col, K, strategy):
    sel, w, _ = _weights(inp, p, ha, K, strategy, False)
    wt = inp["m"][sel] * w / inp["rho"][sel]
    den = np.sum(wt)
    return np.nan if den == 0 else float(np.sum(inp["cols"][col][sel] * wt) / den)


def gradient_quantity(inp, p, ha, col, K, strategy):
    sel, w, g = _weights(inp, p, ha, K, strategy, True)
    mr = inp["m"][sel] / inp["rho"][sel]
    wt = mr * w
    S1 = np.sum(wt)
    if S1 == 0:
        return (np.nan,) * 3
    A = inp["cols"][col][sel]
    Abar = np.sum(A * wt) / S1
    return tuple(float(np.sum(A * mr * gd) - Abar * np.sum(mr * gd)) for gd in g)


def divergence(inp, p, ha, cols, K, strategy):
    sel, w, g = _weights(inp, p, ha, K, strategy, True)
    mr = inp["m"][sel] / inp["rho"][sel]
    wt = mr * w
    S1 = np.sum(wt)
    if S1 == 0:
        return np.nan
    A = [inp["cols"][c][sel] for c in cols]
    Abar = [np.sum(a * wt) / S1 for a in A]
    f = sum(np.sum(a * mr * gd) for a, gd in zip(A, g))
    b = sum(ab * np.sum(mr * gd) for ab, gd in zip(Abar, g))
    return float(f - b)


def curl(inp, p, ha, cols, K, strategy):
    sel, w, g = _weights(inp, p, ha, K, strategy, True)
    mr = inp["m"][sel] / inp["rho"][sel]
    wt = mr * w
    S1 = np.sum(wt)
    if S1 == 0:
        return (np.nan,) * 3
    Ax, Ay, Az = [inp["cols"][c][sel] for c in cols]
    gx, gy, gz = (mr * gd for gd in g)
    fx = np.sum(Ay * gz - Az * gy)
    fy = np.sum(Az * gx - Ax * gz)
    fz = np.sum(Ax * gy - Ay * gx)
    ax, ay, az = (np.sum(a * wt) / S1 for a in (Ax, Ay, Az))
    bx, by, bz = np.sum(gx), np.sum(gy), np.sum(gz)
    return (float(-(fx - (ay * bz - az * by))), float(-(fy - (az * bx - ax * bz))), float(-(fz - (ax * by - ay * bx))))


def nearest_h(inp, p):
    d2 = (p[0] - inp["x"]) ** 2 + (p[1] - inp["y"]) ** 2 + (p[2] - inp["z"]) ** 2
    return float(inp["h"][int(np.argmin(d2))])


def line_d2(inp, o, m):
    D = [inp["x"] - o[0], inp["y"] - o[1], inp["z"] - o[2]]
    D2 = D[0] ** 2 + D[1] ** 2 + D[2] ** 2
    Dm = D[0] * m[0] + D[1] * m[1] + D[2] * m[2]
    return D2 - Dm * Dm


def line_quantity(inp, o, m, col, K, lineW, shepard=True):
    """Scatter line-integrated quantity; lineW(K, r, h) evaluates the tabulated kernel."""
    d2 = line_d2(inp, o, m)
    hb = inp["h"]
    sel = d2 <= (SUPPORT[K] * hb) ** 2
    r = np.sqrt(np.maximum(d2[sel], 0))
    w = np.array([lineW(K, ri, hi) for ri, hi in zip(r, hb[sel])])
    wt = inp["m"][sel] * w / inp["rho"][sel]
    S1 = np.sum(wt)
    if S1 == 0:
        return np.nan
    num = np.sum(inp["cols"][col][sel] * wt)
    return float(num / S1) if shepard else float(num)


def number_density(inp, p, ha, K, strategy):
    """brute_number_density (test/interpolation_test_common.jl:104-118): sum W."""
    sel, w, _ = _weights(inp, p, ha, K, strategy, False)
    return float(np.sum(w))


def gradient_density(inp, p, ha, K, strategy):
    """brute_gradient_density (test/interpolation_test_common.jl:156-186): the
    unnormalised rho = sum m W weights the mean term."""
    sel, w, g = _weights(inp, p, ha, K, strategy, True)
    m, rb = inp["m"][sel], inp["rho"][sel]
    rho = np.sum(m * w)
    if rho == 0:
        return (np.nan,) * 3
    return tuple(float(np.sum(m * gd) - rho * np.sum(m / rb * gd)) for gd in g)


def line_density(inp, o, m, ha, K, lineW, strategy):
    """brute_line_integrated_density (test/interpolation_test_common.jl:139-154):
    sum m_b W along the line with the strategy's radius and kernel."""
    d2 = line_d2(inp, o, m)
    hb = inp["h"]
    Kv = SUPPORT[K]
    rad = Kv * ha * np.ones_like(hb) if strategy == 0 else (Kv * hb if strategy == 1 else Kv * np.maximum(ha, hb))
    sel = d2 <= rad * rad
    r = np.sqrt(np.maximum(d2[sel], 0))
    if strategy == 2:
        w = np.array([0.5 * (lineW(K, ri, ha) + lineW(K, ri, hi)) for ri, hi in zip(r, hb[sel])])
    else:
        w = np.array([lineW(K, ri, ha if strategy == 0 else hi) for ri, hi in zip(r, hb[sel])])
    return float(np.sum(inp["m"][sel] * w)) if w.size else 0.0


def line_quantity_ha(inp, o, m, ha, col, K, lineW, strategy, shepard=True):
    """brute_line_integrated_quantity for any strategy (caller's ha)."""
    d2 = line_d2(inp, o, m)
    hb = inp["h"]
    Kv = SUPPORT[K]
    rad = Kv * ha * np.ones_like(hb) if strategy == 0 else (Kv * hb if strategy == 1 else Kv * np.maximum(ha, hb))
    sel = d2 <= rad * rad
    r = np.sqrt(np.maximum(d2[sel], 0))
    if strategy == 2:
        w = np.array([0.5 * (lineW(K, ri, ha) + lineW(K, ri, hi)) for ri, hi in zip(r, hb[sel])])
    else:
        w = np.array([lineW(K, ri, ha if strategy == 0 else hi) for ri, hi in zip(r, hb[sel])])
    wt = inp["m"][sel] * w / inp["rho"][sel]
    S1 = np.sum(wt)
    if S1 == 0:
        return np.nan
    num = np.sum(inp["cols"][col][sel] * wt)
    return float(num / S1) if shepard else float(num)
