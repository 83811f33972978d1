"""SimGroup-compatible simulation groups on the MI355X core (the caller side
of ``run_simulations``; reference src/cubnm/sim)."""
from .base import WORST_SCORES, MultiSimGroup, SimGroup, create_multi_sim_group  # noqa: F401
from .models import JRSimGroup, KuramotoSimGroup, WCSimGroup, rWWExSimGroup, rWWSimGroup  # noqa: F401
