"""Magnetometric demagnetising factors of a rectangular prism in closed form
(A. Aharoni, J. Appl. Phys. 83, 3432 (1998), eq. 1).

The volume average of the demagnetising field of a uniformly magnetised
prism is exactly -Ms N m with N diagonal; the Newell cell tensor is exact
for cell averages, so the cell-summed field of a fully occupied grid gives
the same N to roundoff at any size.
"""
import math


def _dz(a, b, c):
    """Factor along the axis of half-length c; a, b the other half-lengths."""
    r = math.sqrt(a * a + b * b + c * c)
    ab = math.sqrt(a * a + b * b)
    bc = math.sqrt(b * b + c * c)
    ac = math.sqrt(a * a + c * c)
    s = ((b * b - c * c) / (2 * b * c) * math.log((r - a) / (r + a))
         + (a * a - c * c) / (2 * a * c) * math.log((r - b) / (r + b))
         + b / (2 * c) * math.log((ab + a) / (ab - a))
         + a / (2 * c) * math.log((ab + b) / (ab - b))
         + c / (2 * a) * math.log((bc - b) / (bc + b))
         + c / (2 * b) * math.log((ac - a) / (ac + a))
         + 2 * math.atan(a * b / (c * r))
         + (a ** 3 + b ** 3 - 2 * c ** 3) / (3 * a * b * c)
         + (a * a + b * b - 2 * c * c) / (3 * a * b * c) * r
         + c / (a * b) * (ac + bc)
         - (ab ** 3 + bc ** 3 + ac ** 3) / (3 * a * b * c))
    return s / math.pi


def prism_factors(lx, ly, lz):
    """(Nx, Ny, Nz) of an lx x ly x lz prism (full edge lengths)."""
    a, b, c = lx / 2, ly / 2, lz / 2
    return _dz(b, c, a), _dz(a, c, b), _dz(a, b, c)
