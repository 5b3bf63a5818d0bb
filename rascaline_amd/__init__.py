"""MI355X-native SOAP / LODE descriptor hot path with featomic's calculator API.

The calculators below drive librascaline_amd.so (HIP kernels for gfx950) through
the featomic C API; results are metatensor-style TensorMaps.
"""

from ._c_lib import FeatomicError  # noqa: F401
from .calculator_base import CalculatorBase, neighbors  # noqa: F401
from .calculators import (LodeSphericalExpansion, NeighborList, SoapPowerSpectrum,  # noqa: F401
                          SoapRadialSpectrum, SphericalExpansion, SphericalExpansionByPair)
from .systems import System  # noqa: F401
from .tensormap import Labels, TensorBlock, TensorMap  # noqa: F401
