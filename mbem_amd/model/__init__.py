"""Region model, unknown layout and block equations (mbem's src/mbem/model)."""
from .core import (FAULT_KINDS, FAULT_ORIENTATION, BCType, Patch, Region,  # noqa: F401
                   RegionModel)
from .equations import (BlockSystem, BlockTerm, RhsTerm, RhsValue,  # noqa: F401
                        generate_system)
from .layout import Slot, UnknownLayout  # noqa: F401
