"""Exception types of the reference (Julia) mirrored in Python.

The C ABI returns negative status codes; `raise_for_status` maps each to the
exception the reference throws in the same situation (see the PH_ERR_* comments
in include/partia_hip.h).
"""


class ArgumentError(ValueError):
    """Julia's ArgumentError."""


class DimensionMismatch(ValueError):
    """Julia's DimensionMismatch."""


class MethodError(TypeError):
    """Julia's MethodError: no method matches the argument types."""


class JuliaAssertionError(AssertionError):
    """Julia's AssertionError (an internal invariant failed)."""


class DeviceError(RuntimeError):
    """HIP runtime failure, device out of memory, or no usable GPU."""


def raise_for_status(code, message):
    if code == -1 or code == -3 or code == -4 or code == -5:
        raise ArgumentError(message)
    if code == -2:
        raise DimensionMismatch(message)
    if code == -8:
        raise JuliaAssertionError(message)
    raise DeviceError("status %d: %s" % (code, message))
