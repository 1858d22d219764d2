"""SPH kernel type tags, interpolation strategies and execution backends.

Mirrors KI/kernel_function/kernels/*.jl (support radius, normalisation and
typical neighbour count of M4/M5/M6 splines and C2/C4/C6 Wendland kernels),
KI/interpolation_setup/InterpolationStrategy.jl:1-5 and
KI/ExecutionBackend/AbstractExecutionBackend.jl:1-2 + ext backends.
The kernel shapes themselves are evaluated only on the GPU.
"""
import math

import numpy as np


class AbstractSPHKernel:
    id = -1
    support = 0.0
    nneigh = 0
    # (1D, 2D, 3D) normalisation as written in the reference, as (numerator, denominator, times_pi)
    _norm = ((0, 1, False), (0, 1, False), (0, 1, False))


def _norm_value(spec, dtype):
    T = np.dtype(dtype).type
    num, den, with_pi = spec
    if with_pi:
        return T(num) / (T(den) * T(math.pi))
    return T(num) / T(den)


class M4_spline(AbstractSPHKernel):
    id, support, nneigh = 0, 2.0, 57
    _norm = ((4, 3, False), (10, 7, True), (1, 1, True))


class M5_spline(AbstractSPHKernel):
    id, support, nneigh = 1, 2.5, 113
    _norm = ((1, 24, False), (96, 1199, True), (0.05, 1, True))


class M6_spline(AbstractSPHKernel):
    id, support, nneigh = 2, 3.0, 112
    _norm = ((1, 120, False), (7, 478, True), (1, 120, True))


class C2_Wendland(AbstractSPHKernel):
    id, support, nneigh = 3, 2.0, 92
    _norm = ((5, 8, False), (7, 4, True), (21, 16, True))


class C4_Wendland(AbstractSPHKernel):
    id, support, nneigh = 4, 2.0, 137
    _norm = ((3, 4, False), (9, 4, True), (495, 256, True))


class C6_Wendland(AbstractSPHKernel):
    id, support, nneigh = 5, 2.0, 356
    _norm = ((64, 55, False), (78, 28, True), (1365, 512, True))


ALL_KERNELS = (M4_spline, M5_spline, M6_spline, C2_Wendland, C4_Wendland, C6_Wendland)


def KernelFunctionValid(K, dtype=np.float64):
    """Support radius in units of h (M4_spline.jl:66 and siblings)."""
    return np.dtype(dtype).type(K.support)


def KernelFunctionnorm(K, D, dtype=np.float64):
    """Normalisation sigma_D (M4_spline.jl:99-101 and siblings)."""
    if D not in (1, 2, 3):
        raise ValueError("dimension must be 1, 2 or 3")
    return _norm_value(K._norm[D - 1], dtype)


def KernelFunctionNneigh(K):
    """Typical neighbour count (M4_spline.jl:127 and siblings)."""
    return K.nneigh


class AbstractInterpolationStrategy:
    id = -1


class itpGather(AbstractInterpolationStrategy):
    """Smoothing length of the sample (nearest particle's h)."""
    id = 0


class itpScatter(AbstractInterpolationStrategy):
    """Smoothing length of each neighbour."""
    id = 1


class itpSymmetric(AbstractInterpolationStrategy):
    """Average of the gather and scatter kernels."""
    id = 2


class AbstractExecutionBackend:
    pass


class CPUComputeBackend(AbstractExecutionBackend):
    """The reference's own CPU path.  This package implements no CPU method:
    every driver called with this backend raises MethodError."""


class HIPComputeBackend(AbstractExecutionBackend):
    """MI355X backend: one HIP context (GPU + stream) per instance."""

    def __init__(self, device=0):
        self.device = device
        self._ctx = None

    @property
    def context(self):
        if self._ctx is None:
            from ._native import Context
            self._ctx = Context(self.device)
        return self._ctx

    def __repr__(self):
        return "HIPComputeBackend(device=%d)" % self.device


_default_backend = None


def default_backend():
    global _default_backend
    if _default_backend is None:
        _default_backend = HIPComputeBackend(0)
    return _default_backend
