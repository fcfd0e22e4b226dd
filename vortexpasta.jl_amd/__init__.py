"""vortexpasta.jl_amd — MI355X-native Ewald-split Biot–Savart evaluation.

Host-side mirror of VortexPasta.jl's BiotSavart interface over the libvpbs C
ABI (include/vpbs.h). The directory name contains a dot, so it is loaded with
``__graft_entry__.load_package()`` (module name ``vortexpasta_jl_amd``).
"""

from . import diagnostics, filaments, forcing, learned, params, pointdata, tangles, timestepping  # noqa: F401
from .biotsavart import (  # noqa: F401
    BiotSavartCache,
    background_vorticity_correction,
    compute_on_nodes_,
    get_longrange_field_fourier,
    init_cache,
    velocity_on_nodes,
    velocity_on_nodes_,
)
from .filaments import ClosedFilament, FilamentBatch, PeriodicLine, Ring, TrefoilKnot, define_curve  # noqa: F401
from .params import (  # noqa: F401
    GaussianSplitting,
    GaussLegendre,
    HIPLongRangeBackend,
    HIPShortRangeBackend,
    NoQuadrature,
    ParamsBiotSavart,
)
from .pointdata import PointData, add_point_charges  # noqa: F401
