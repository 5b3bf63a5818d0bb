"""The SOAP, LODE and neighbour-list calculators (python/featomic/featomic/calculators.py:44-307)."""

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


class SphericalExpansionByPair(CalculatorBase):
    """The spherical expansion split per directed pair (i, j, cell shift): the
    contribution of atom j to the density around atom i, keys (o3_lambda, o3_sigma,
    first_atom_type, second_atom_type), self pairs in the l = 0 blocks
    (featomic/src/calculators/soap/spherical_expansion_pair.rs:547-840). Summing the
    samples over the second atom gives the spherical expansion."""

    def __init__(self, *, cutoff, density, basis):
        parameters = {"cutoff": cutoff, "density": density, "basis": basis}
        super().__init__("spherical_expansion_by_pair", json.dumps(_to_json(parameters)))


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


class NeighborList(CalculatorBase):
    """Pair vectors of the neighbour list within ``cutoff``, one block per (first,
    second) atom-type pair (featomic/src/calculators/neighbor_list.rs; the Python
    class of python/featomic/featomic/calculators.py:44-74). A half list stores each
    pair once, from the lower type to the higher; a full list stores both
    directions; ``self_pairs`` adds (i, i) pairs with a zero vector. Samples are
    (system, first_atom, second_atom, cell_shift_a/b/c), one component ``pair_xyz``
    and one property ``distance`` = 1. The pairs come from the GPU cell list."""

    def __init__(self, *, cutoff, full_neighbor_list, self_pairs=False):
        parameters = {"cutoff": cutoff, "full_neighbor_list": full_neighbor_list, "self_pairs": self_pairs}
        super().__init__("neighbor_list", json.dumps(_to_json(parameters)))
