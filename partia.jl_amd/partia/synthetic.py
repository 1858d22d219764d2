"""Seeded synthetic particle sets and sample layouts for the benchmark workloads.

Random numbers come from a counter-based SplitMix64: the k-th 64-bit output of
stream `seed` is mix(seed + k * 0x9E3779B97F4A7C15), so any slice of a stream is
reproducible without generating what precedes it.  Doubles are (u >> 11) * 2^-53
and normals use Box-Muller.

Workloads (sizes as in the benchmark configurations):
  uniform_box      N particles uniform in [0,1]^3, equal masses, rho = 1,
                   h = eta (m/rho)^(1/3), A = sin(2 pi x) cos(2 pi y)
  plummer_sphere   Plummer sphere (scale radius a, truncated at 10 a), m = 1/N,
                   h = eta (m / rho_Plummer(r))^(1/3), v = solid-body rotation
                   about z plus N(0, sigma_v^2) noise, A as above
  uniform_plus_clump  half uniform in [0,1]^3, half a Plummer clump (a = 0.05,
                   truncated at 10 a) at the centre; m = 1/N, h from the analytic
                   density of both components, v = rotation about the clump axis
  request_mix      sampling requests of uneven size over [0,1]^3: point lists
                   (count log-uniform), oriented planes (n^2) and boxes (n^3)
  lognormal_field  particles drawn with probability proportional to a log-normal
                   density exp(delta - s^2/2), delta a Gaussian random field with
                   P(k) ~ k^(-11/3) on a periodic base grid
"""
import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, start, count):
    """Outputs start .. start+count-1 of SplitMix64 stream `seed` (uint64 array)."""
    k = np.arange(start + 1, start + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + k * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


class Stream:
    """Sequential draws from one SplitMix64 stream."""

    def __init__(self, seed):
        self.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self.pos = 0

    def uniform(self, count):
        u = splitmix64(self.seed, self.pos, count)
        self.pos += count
        return (u >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, count):
        u1 = self.uniform(count)
        u2 = self.uniform(count)
        u1 = np.where(u1 <= 0.0, 1.0 / 9007199254740992.0, u1)
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def scalar_field(x, y):
    return np.sin(2.0 * np.pi * x) * np.cos(2.0 * np.pi * y)


def uniform_box(n, seed=1, mass=1e-5, eta=1.2):
    st = Stream(seed)
    x, y, z = st.uniform(n), st.uniform(n), st.uniform(n)
    m = np.full(n, mass)
    rho = np.ones(n)
    h = np.full(n, eta * (mass / 1.0) ** (1.0 / 3.0))
    return dict(x=x, y=y, z=z, m=m, h=h, rho=rho, A=scalar_field(x, y))


def plummer_density(r, a=0.1, M=1.0):
    return 3.0 * M / (4.0 * np.pi * a ** 3) * (1.0 + (r / a) ** 2) ** (-2.5)


def plummer_sphere(n, seed=2, a=0.1, rmax_over_a=10.0, eta=1.2, sigma_v=0.01):
    st = Stream(seed)
    mfrac = rmax_over_a ** 3 / (1.0 + rmax_over_a ** 2) ** 1.5  # enclosed mass fraction at the cut
    u = st.uniform(n) * mfrac
    u = np.where(u <= 0.0, 1e-300, u)
    r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    cost = 2.0 * st.uniform(n) - 1.0
    phi = 2.0 * np.pi * st.uniform(n)
    sint = np.sqrt(np.maximum(0.0, 1.0 - cost * cost))
    x, y, z = r * sint * np.cos(phi), r * sint * np.sin(phi), r * cost
    m = np.full(n, 1.0 / n)
    rho = plummer_density(r, a)
    h = eta * (m / rho) ** (1.0 / 3.0)
    vx = -y + sigma_v * st.normal(n)
    vy = x + sigma_v * st.normal(n)
    vz = sigma_v * st.normal(n)
    return dict(x=x, y=y, z=z, m=m, h=h, rho=rho, A=scalar_field(x, y), vx=vx, vy=vy, vz=vz)


def uniform_plus_clump(n, seed=4, a=0.05, rmax_over_a=10.0, eta=1.2, sigma_v=0.01):
    st = Stream(seed)
    nu = n // 2
    nc = n - nu
    xu, yu, zu = st.uniform(nu), st.uniform(nu), st.uniform(nu)
    mfrac = rmax_over_a ** 3 / (1.0 + rmax_over_a ** 2) ** 1.5
    u = st.uniform(nc) * mfrac
    u = np.where(u <= 0.0, 1e-300, u)
    r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    cost = 2.0 * st.uniform(nc) - 1.0
    phi = 2.0 * np.pi * st.uniform(nc)
    sint = np.sqrt(np.maximum(0.0, 1.0 - cost * cost))
    x = np.concatenate([xu, 0.5 + r * sint * np.cos(phi)])
    y = np.concatenate([yu, 0.5 + r * sint * np.sin(phi)])
    z = np.concatenate([zu, 0.5 + r * cost])
    m = np.full(n, 1.0 / n)
    rc = np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2)
    rho = nu / n + plummer_density(rc, a, M=nc / n / mfrac) * (rc <= rmax_over_a * a)
    h = eta * (m / rho) ** (1.0 / 3.0)
    vx = -(y - 0.5) + sigma_v * st.normal(n)
    vy = (x - 0.5) + sigma_v * st.normal(n)
    vz = sigma_v * st.normal(n)
    return dict(x=x, y=y, z=z, m=m, h=h, rho=rho, A=scalar_field(x, y), vx=vx, vy=vy, vz=vz)


def random_rotation(st):
    """Rotation matrix of a uniformly random unit quaternion."""
    q = st.normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def request_mix(nreq, seed=4, points=(1e3, 1e6), plane=(256, 1024), box=(32, 128), half_width=0.3):
    """Request descriptions, kind round-robin over point list / plane / box.
    Point lists: count log-uniform in `points`, positions uniform in [0,1]^3.
    Planes (n x n) and boxes (n^3): n log-uniform in `plane` / `box`, a random
    frame (right, up, forward = columns 0, 2, 1 of a random rotation) at a
    uniform position in [0.2, 0.8]^3, local axes (-w, w, n)."""
    st = Stream(seed + 77)
    out = []
    for r in range(nreq):
        kind = ("points", "plane", "box")[r % 3]
        u = st.uniform(1)[0]
        if kind == "points":
            lo, hi = np.log(points[0]), np.log(points[1])
            count = int(np.exp(lo + u * (hi - lo)))
            out.append(dict(kind=kind, count=count, seed=int(st.uniform(1)[0] * 2 ** 31)))
        else:
            rng_n = plane if kind == "plane" else box
            lo, hi = np.log(rng_n[0]), np.log(rng_n[1])
            nn = int(np.exp(lo + u * (hi - lo)))
            R = random_rotation(st)
            pos = 0.2 + 0.6 * st.uniform(3)
            axis = (-half_width, half_width, nn)
            out.append(dict(kind=kind, n=nn, position=pos, right=R[:, 0], up=R[:, 2], forward=R[:, 1], axis=axis))
    return out


def gaussian_random_field(ng, seed, slope=-11.0 / 3.0, sigma_ln=1.5):
    """Periodic Gaussian random field on ng^3 with P(k) ~ k^slope, unit-normalised to sigma_ln."""
    st = Stream(seed)
    white = st.normal(ng ** 3).reshape(ng, ng, ng)
    fk = np.fft.rfftn(white)
    kx = np.fft.fftfreq(ng)[:, None, None]
    ky = np.fft.fftfreq(ng)[None, :, None]
    kz = np.fft.rfftfreq(ng)[None, None, :]
    k = np.sqrt(kx * kx + ky * ky + kz * kz)
    k[0, 0, 0] = 1.0
    amp = k ** (slope / 2.0)
    amp[0, 0, 0] = 0.0
    delta = np.fft.irfftn(fk * amp, s=(ng, ng, ng), axes=(0, 1, 2))
    delta *= sigma_ln / delta.std()
    return delta


def lognormal_field(n, seed=3, ng=128, sigma_ln=1.5, eta=1.2):
    delta = gaussian_random_field(ng, seed, sigma_ln=sigma_ln)
    rho_cell = np.exp(delta - 0.5 * sigma_ln ** 2)
    p = rho_cell.ravel() / rho_cell.sum()
    cdf = np.cumsum(p)
    st = Stream(seed + 1000)
    cell = np.searchsorted(cdf, st.uniform(n) * cdf[-1], side="right")
    cell = np.minimum(cell, ng ** 3 - 1)
    ix, iy, iz = np.unravel_index(cell, (ng, ng, ng))
    x = (ix + st.uniform(n)) / ng
    y = (iy + st.uniform(n)) / ng
    z = (iz + st.uniform(n)) / ng
    m = np.full(n, 1.0 / n)
    rho = rho_cell.ravel()[cell] / rho_cell.mean()
    h = eta * (m / rho) ** (1.0 / 3.0)
    return dict(x=x, y=y, z=z, m=m, h=h, rho=rho, A=scalar_field(x, y))


def linrange(lo, hi, n, dtype=np.float64):
    """collect(LinRange{T}(T(lo), T(hi), n)) (Grids/grids/StructuredGrid.jl:288).

    Element j (0-based) is T((1 - t)*a + t*b) with t = j/d, d = max(n-1, 1), a and
    b the endpoints converted to T; t is a Float64 quotient of integers, so the
    lerp is evaluated in Float64 (no fused multiply-add) and rounded to T once.
    """
    T = np.dtype(dtype).type
    a, b = np.float64(T(lo)), np.float64(T(hi))
    n = int(n)
    if n < 0:
        raise ValueError("negative length")
    d = float(max(n - 1, 1))
    t = np.arange(n, dtype=np.float64) / d
    return ((1.0 - t) * a + t * b).astype(dtype)
