"""The SOAP and LODE calculators (python/featomic/featomic/calculators.py:120-307)."""

import json

from .calculator_base import CalculatorBase


def _to_json(value):
    """Hyper-parameters as plain JSON types (numpy scalars included)."""
    if isinstance(value, dict):
        return {str(k): _to_json(v) for k, v in value.items()}
    if isinstance(value, (list, tuple)):
        return [_to_json(v) for v in value]
    if hasattr(value, "item"):
        return value.item()
    return value


class SphericalExpansion(CalculatorBase):
    """SOAP spherical expansion coefficients <n l m | rho_i^{a}>."""

    def __init__(self, *, cutoff, density, basis):
        parameters = {"cutoff": cutoff, "density": density, "basis": basis}
        super().__init__("spherical_expansion", json.dumps(_to_json(parameters)))


class SoapPowerSpectrum(CalculatorBase):
    """SOAP power spectrum, sum_m <n1 l m|rho^a1> <n2 l m|rho^a2>."""

    def __init__(self, *, cutoff, density, basis):
        parameters = {"cutoff": cutoff, "density": density, "basis": basis}
        super().__init__("soap_power_spectrum", json.dumps(_to_json(parameters)))


class SoapRadialSpectrum(CalculatorBase):
    """SOAP radial spectrum <n | rho_i^{a}> (the l = 0 spherical expansion), with the
    reference's parameters: cutoff, density and basis = {radial, spline_accuracy}."""

    def __init__(self, *, cutoff, density, basis):
        parameters = {"cutoff": cutoff, "density": density, "basis": basis}
        super().__init__("soap_radial_spectrum", json.dumps(_to_json(parameters)))


class LodeSphericalExpansion(CalculatorBase):
    """Long-distance equivariant (LODE) spherical expansion, computed in k-space."""

    def __init__(self, *, density, basis, k_cutoff=None):
        parameters = {"k_cutoff": k_cutoff, "density": density, "basis": basis}
        super().__init__("lode_spherical_expansion", json.dumps(_to_json(parameters)))
