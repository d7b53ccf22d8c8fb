"""Python entry points for callers of the reference's PyInteractor.

``TorchProjector`` applies the forward projector and its exact adjoint to
device tensors as a torch.autograd.Function, configured from the same
header objects the reference's PyInteractor passes to FIR_ForwardProj and
FIR_BackProj.
"""
from .projector import (TorchProjector, proj_spec, scanner_spec,  # noqa: F401
                        vol_spec)
