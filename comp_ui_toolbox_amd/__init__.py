"""MI355X-native Born scattering operator pair (forward and adjoint) of the
Computational Ultrasound Imaging Toolbox, behind the C ABI in include/born.h."""
from .configs import SequenceSpec  # noqa: F401
from ._lib import BornError, BornLibraryError  # noqa: F401
from .operator_born import OperatorBorn, OperatorBatch, combined_quick_gpu  # noqa: F401
