"""InterpolationCatalog: which columns to interpolate and in which output order.

Mirrors KI/interpolation_setup/InterpolationCatalog.jl: name -> column-slot
resolution (1-based slots), vector names expanded with x/y/z suffixes
(:191-197), output order scalars, gradients (x, y, z), divergences, curls
(x, y, z) (:199-220), Shepard normalisation for every scalar except rho/ρ
(:222-225), and the concise catalog handed to the kernels (:26-32, 268-276).
Names are Python strings (Julia Symbols).
"""
from .errors import ArgumentError, DimensionMismatch

_XYZ = ("x", "y", "z")
_DERIV = ("ˣ", "ʸ", "ᶻ")
_NO_SHEPARD = ("ρ", "rho")


def _vector_components(name, D=3):
    return tuple(name + _XYZ[i] for i in range(D))


def _ordered_quantity_names(D, scalars, grads, divs, curls):
    names = list(scalars)
    for base in grads:
        names += ["∇" + base + _DERIV[i] for i in range(D)]
    for base in divs:
        names.append("∇⋅" + base)
    for base in curls:
        names += ["∇×" + base + _DERIV[i] for i in range(D)]
    return tuple(names)


def _shepard_normalization_flag(scalar_names):
    return tuple(name not in _NO_SHEPARD for name in scalar_names)


class InterpolationCatalogConcise:
    __slots__ = ("D", "scalar_slots", "scalar_snormalization", "grad_slots", "div_slots", "curl_slots")

    def __init__(self, D, scalar_slots, scalar_snormalization, grad_slots, div_slots, curl_slots):
        self.D = D
        self.scalar_slots = tuple(scalar_slots)
        self.scalar_snormalization = tuple(scalar_snormalization)
        self.grad_slots = tuple(grad_slots)
        self.div_slots = tuple(tuple(s) for s in div_slots)
        self.curl_slots = tuple(tuple(s) for s in curl_slots)

    @property
    def L(self):
        return len(self.scalar_slots) + 3 * len(self.grad_slots) + len(self.div_slots) + 3 * len(self.curl_slots)

    def native(self):
        """ph_catalog arguments for the C ABI."""
        from ._native import CatalogArgs
        return CatalogArgs(sslots=self.scalar_slots, sflags=[1 if f else 0 for f in self.scalar_snormalization],
                           gslots=self.grad_slots, dslots=[s for t in self.div_slots for s in t],
                           cslots=[s for t in self.curl_slots for s in t])


class InterpolationCatalog:
    """InterpolationCatalog{D, N, G, Div, C, L}."""

    def __init__(self, D=3, scalar_names=(), scalar_slots=(), grad_names=(), grad_slots=(), div_names=(),
                 div_slots=(), curl_names=(), curl_slots=()):
        # lower constructor (InterpolationCatalog.jl:71-102)
        N, G, Div, C = len(scalar_names), len(grad_names), len(div_names), len(curl_names)
        if len(scalar_slots) != N:
            raise DimensionMismatch("scalar_slots length %d != scalar_names length %d" % (len(scalar_slots), N))
        if len(grad_slots) != G:
            raise DimensionMismatch("grad_slots length %d != grad_names length %d" % (len(grad_slots), G))
        if len(div_slots) != Div:
            raise DimensionMismatch("div_slots length %d != div_names length %d" % (len(div_slots), Div))
        if len(curl_slots) != C:
            raise DimensionMismatch("curl_slots length %d != curl_names length %d" % (len(curl_slots), C))
        self.D = D
        self.scalar_names = tuple(scalar_names)
        self.scalar_slots = tuple(int(s) for s in scalar_slots)
        self.scalar_snormalization = _shepard_normalization_flag(self.scalar_names)
        self.grad_names = tuple(grad_names)
        self.grad_slots = tuple(int(s) for s in grad_slots)
        self.div_names = tuple(div_names)
        self.div_slots = tuple(tuple(int(x) for x in s) for s in div_slots)
        self.curl_names = tuple(curl_names)
        self.curl_slots = tuple(tuple(int(x) for x in s) for s in curl_slots)
        self.ordered_names = _ordered_quantity_names(D, self.scalar_names, self.grad_names, self.div_names,
                                                     self.curl_names)

    @classmethod
    def from_columns(cls, column_names, D=3, scalars=(), gradients=(), divergences=(), curls=()):
        """InterpolationCatalog(column_names, Val(D); scalars, gradients, divergences, curls)
        (InterpolationCatalog.jl:141-188): KeyError on unknown names."""
        column_names = tuple(column_names)
        if len(set(column_names)) != len(column_names):
            raise ArgumentError("column_names must be unique")
        index = {name: i + 1 for i, name in enumerate(column_names)}
        scalar_slots = tuple(index[s] for s in scalars)
        grad_slots = tuple(index[g] for g in gradients)
        div_slots = tuple(tuple(index[c] for c in _vector_components(v, D)) for v in divergences)
        curl_slots = tuple(tuple(index[c] for c in _vector_components(v, D)) for v in curls)
        return cls(D, scalars, scalar_slots, gradients, grad_slots, divergences, div_slots, curls, curl_slots)

    @property
    def L(self):
        return len(self.ordered_names)

    @property
    def shape(self):
        """(N, G, Div, C)."""
        return (len(self.scalar_names), len(self.grad_names), len(self.div_names), len(self.curl_names))


def to_concise_catalog(cat):
    return InterpolationCatalogConcise(cat.D, cat.scalar_slots, cat.scalar_snormalization, cat.grad_slots,
                                       cat.div_slots, cat.curl_slots)


def scalar_index(cat, name):
    return cat.scalar_slots[cat.scalar_names.index(name)]


def grad_slot(cat, name):
    return cat.grad_slots[cat.grad_names.index(name)]


def div_slots(cat, name):
    return cat.div_slots[cat.div_names.index(name)]


def curl_slots(cat, name):
    return cat.curl_slots[cat.curl_names.index(name)]


def ordered_quantity_names(cat):
    return cat.ordered_names
