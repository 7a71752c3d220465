"""taichi_lbm on MI355X: the reference's D2Q9 API (Chiuzzmm/taichi-lbm,
taichi_lbm/__init__.py:4-22) over hand-written HIP kernels for gfx950.

Import this package as ``taichi_lbm`` (put its parent directory,
``taichi-lbm_amd/``, on ``sys.path``).  Names outside the hot path are
exported so that scripts import unchanged; constructing them raises
NotImplementedError.
"""
from .lbm import LBField, MacroscopicEngine
from .BoundaryCondition import (BoundarySpec, BoundaryClassifier, MaskAndGroup, VelocityBB, VelocityEquilibrium,
                                VelocityNEEM, PressureABC, PressureNEEM, InsideBoundary, BoundaryEngine,
                                BounceBackWall, FluidBoundary, PeriodicAllBoundary, OpenNeumann,
                                OpenExtrapolation, OpenConvective1order, OpenConvective2order)
from .CollisionEngine import BGKCollision, TRTCollision, MRTCollision, HuangMRTCollision
from .ibm import IBMNodes, Sphere, SphereIB, IBDEMCouplerEngine, IBEngine
from .ExternalForce import ShanChenForceC1, ShanChenForceC2
from .SCBoundaryCondition import (BoundaryEngineSC, OpenNeumannPsi, OpenConvective1orderPsi,
                                  OpenConvective2orderPsi, OpenExtrapolationPsi, PeriodicPsi_LR, PeriodicPsi_BT,
                                  FixedPsi)
from .EOS import SC_psi, vdW_psi, RK_psi, PR_psi, CS_psi
from .FluidModels import PowerLawFluid, NewtonianFluid
from .PostProcessing import PostProcessingEngine
from .stepper import Stepper
from ._codes import UnsupportedBoundaryList
from ._hip import LBMError

__all__ = [n for n in dir() if not n.startswith("_")]
