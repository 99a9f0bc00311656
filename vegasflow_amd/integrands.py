"""Built-in synthetic integrand families.

An instance is a plain callable ``f(xarr, weight=None)`` on a torch tensor
(n_events, n_dim), so it works everywhere a user integrand does (the
unfused sample -> f(x) -> digest path). When one is handed to ``compile`` and
n_dim <= 16, the integrators instead evaluate it in registers inside the fused
kernels (family id plus parameter block, ``mv_integrand`` in
include/mivegas.h); results agree with the torch form to rounding.

Families (all on the unit hypercube unless integration limits are given):

- ``Gaussian(a, centre)``: (1/(a sqrt(pi)))^d exp(-sum_j ((x_j - c_j)/a)^2),
  the reference's test integrand (test_algs.py:27-36) without its
  zero-valued ``range`` term; integral prod_j (erf((1-c_j)/a) + erf(c_j/a))/2.
- ``Product()``: prod_j x_j (test_misc.py:46-48); integral 2^-d.
- ``GaussianMix(weights, widths, centres)``: sum_k w_k Gaussian(a_k, c_k).
- ``BreitWignerPower(mass, width, power)``: (G/2pi)/((x_0-M)^2 + G^2/4) *
  x_1^p with G the full width; other dimensions flat.
- ``Constant(c)``.
"""
import math

import torch

from . import _abi


def _centre_list(centre, n_dim):
    if isinstance(centre, (int, float)):
        return [float(centre)] * n_dim
    c = [float(v) for v in centre]
    if len(c) != n_dim:
        raise ValueError(f"centre has {len(c)} entries, expected {n_dim}")
    return c


class BuiltinIntegrand:
    """Base class of the fused families."""

    family = None
    __name__ = "builtin"

    def fused_ok(self, n_dim):
        return n_dim <= _abi.FUSED_MAX_DIM

    def to_struct(self, n_dim):
        raise NotImplementedError

    def exact(self, n_dim):
        """Closed-form integral over [0, 1]^n_dim (None when unknown)."""
        return None


class Constant(BuiltinIntegrand):
    family = _abi.FN_CONSTANT
    __name__ = "constant"

    def __init__(self, value=1.0):
        self.value = float(value)

    def fused_ok(self, n_dim):
        return True

    def __call__(self, xarr, weight=None):
        return torch.full((xarr.shape[0],), self.value, dtype=xarr.dtype, device=xarr.device)

    def to_struct(self, n_dim):
        s = _abi.Integrand()
        s.family = self.family
        s.n_comp = 1
        s.comp[0][0] = self.value
        return s

    def exact(self, n_dim):
        return self.value


class Gaussian(BuiltinIntegrand):
    family = _abi.FN_GAUSSIAN
    __name__ = "gaussian"

    def __init__(self, a=0.1, centre=0.5):
        self.a = float(a)
        self.centre = centre

    def prefactor(self, n_dim):
        return (1.0 / (self.a * math.sqrt(math.pi))) ** n_dim

    def __call__(self, xarr, weight=None):
        d = xarr.shape[-1]
        c = torch.as_tensor(_centre_list(self.centre, d), dtype=xarr.dtype, device=xarr.device)
        coef = torch.sum(torch.square((xarr - c) / self.a), dim=1)
        return self.prefactor(d) * torch.exp(-coef)

    def to_struct(self, n_dim):
        s = _abi.Integrand()
        s.family = self.family
        s.n_comp = 1
        s.comp[0][0] = 1.0
        s.comp[0][1] = 1.0 / self.a
        s.comp[0][2] = self.prefactor(n_dim)
        for j, c in enumerate(_centre_list(self.centre, n_dim)):
            s.centre[0][j] = c
        return s

    def exact(self, n_dim):
        out = 1.0
        for c in _centre_list(self.centre, n_dim):
            out *= 0.5 * (math.erf((1.0 - c) / self.a) + math.erf(c / self.a))
        return out


class Product(BuiltinIntegrand):
    family = _abi.FN_PRODUCT
    __name__ = "product"

    def fused_ok(self, n_dim):
        return True

    def __call__(self, xarr, weight=None):
        return torch.prod(xarr, dim=1)

    def to_struct(self, n_dim):
        s = _abi.Integrand()
        s.family = self.family
        s.n_comp = 1
        return s

    def exact(self, n_dim):
        return 0.5 ** n_dim


class GaussianMix(BuiltinIntegrand):
    family = _abi.FN_GAUSSMIX
    __name__ = "gaussian_mix"

    def __init__(self, weights=(0.5, 0.5), widths=(0.05, 0.1), centres=(0.33, 0.67)):
        if not (len(weights) == len(widths) == len(centres)):
            raise ValueError("weights, widths and centres must have the same length")
        if not 1 <= len(weights) <= _abi.MAX_COMP:
            raise ValueError(f"a mixture has 1..{_abi.MAX_COMP} components")
        self.parts = [Gaussian(a, c) for a, c in zip(widths, centres)]
        self.weights = [float(w) for w in weights]

    def __call__(self, xarr, weight=None):
        out = None
        for wk, g in zip(self.weights, self.parts):
            t = wk * g(xarr)
            out = t if out is None else out + t
        return out

    def to_struct(self, n_dim):
        s = _abi.Integrand()
        s.family = self.family
        s.n_comp = len(self.parts)
        for k, (wk, g) in enumerate(zip(self.weights, self.parts)):
            s.comp[k][0] = wk
            s.comp[k][1] = 1.0 / g.a
            s.comp[k][2] = g.prefactor(n_dim)
            for j, c in enumerate(_centre_list(g.centre, n_dim)):
                s.centre[k][j] = c
        return s

    def exact(self, n_dim):
        return sum(wk * g.exact(n_dim) for wk, g in zip(self.weights, self.parts))


class BreitWignerPower(BuiltinIntegrand):
    family = _abi.FN_BWPOWER
    __name__ = "breit_wigner_power"

    def __init__(self, mass=0.5, width=0.02, power=-0.7):
        self.mass = float(mass)
        self.width = float(width)
        self.power = float(power)

    def fused_ok(self, n_dim):
        return n_dim >= 2

    def __call__(self, xarr, weight=None):
        g2 = 0.5 * self.width
        bw = (g2 / math.pi) / (torch.square(xarr[:, 0] - self.mass) + g2 * g2)
        return bw * torch.pow(xarr[:, 1], self.power)

    def to_struct(self, n_dim):
        if n_dim < 2:
            raise ValueError("BreitWignerPower needs n_dim >= 2")
        g2 = 0.5 * self.width
        s = _abi.Integrand()
        s.family = self.family
        s.n_comp = 1
        s.comp[0][0] = self.mass
        s.comp[0][1] = g2
        s.comp[0][2] = g2 / math.pi
        s.comp[0][3] = self.power
        return s

    def exact(self, n_dim):
        g2 = 0.5 * self.width
        bw = (math.atan((1.0 - self.mass) / g2) - math.atan(-self.mass / g2)) / math.pi
        return bw / (self.power + 1.0)
