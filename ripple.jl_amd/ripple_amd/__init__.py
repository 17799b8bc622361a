"""MI355X wave-action time step with the Ripple.jl host surface.

The reference API names map as: SpectralWaveModel, set! -> set_,
time_step! -> time_step_, Simulation, run! -> run_, interior, m0,
root_mean_square_wavenumber, mean_direction, significant_wave_height,
total_action, total_deep_water_energy.  Compute runs in the gfx950 library
libripple_mi355x.so through its C ABI.
"""
from ._lib import RippleError, RippleArgumentError, load as load_library, LIB_PATH
from .grids import (RectilinearGrid, PolarWaveVectorGrid, FrequencyDirectionGrid, centers_to_faces,
                    spectral_cell_measures, xnodes, ynodes, znodes, xfaces, yfaces, zfaces,
                    xspacings, yspacings, zspacings, PERIODIC, BOUNDED, FLAT)
from .model import (SpectralWaveModel, InfiniteDepth, HorizontalField, Clock, Simulation, WENO, set_, time_step_, run_,
                    interior, PseudomomentumVelocities, SpatialAveraging, automatic_q_grid, update_coupling_,
                    compute_pseudomomentum_cell_averages, compute_pseudomomentum_cell_integrals,
                    pseudomomentum_fields, compute_pseudomomentum_tendency_cell_averages,
                    pseudomomentum_tendency_fields, cwcm_momentum_tendency_fields)
from .sources import (SymmetricQuadruplet, DiscreteInteractionApproximation, QuadrupletTransferInteraction,
                      ListedQuadruplets,
                      BulkWindDrag, LinearStormTrack, HollandHurricaneWind, PressureCorrelationInput,
                      MeanSpectrumWhitecapping, LocalSaturationDissipation, WhitecappingDissipation,
                      PrecomputedSources, wind_state)
from .surrogate import (SourceNetwork, SurrogateConfig, training_pairs, attach_surrogate, surrogate_stats,
                        train as train_surrogate)
from .ensemble import Ensemble
from . import validation
from .output import JLD2Writer, IterationInterval, TimeInterval, fetch_output
from .diagnostics import (m0, root_mean_square_wavenumber, mean_direction, significant_wave_height,
                          total_action, total_deep_water_energy)

__all__ = [
    "RippleError", "RippleArgumentError", "load_library", "LIB_PATH",
    "RectilinearGrid", "PolarWaveVectorGrid", "FrequencyDirectionGrid", "centers_to_faces",
    "spectral_cell_measures", "xnodes", "ynodes", "znodes", "xfaces", "yfaces", "zfaces",
    "xspacings", "yspacings", "zspacings", "PERIODIC", "BOUNDED", "FLAT",
    "SpectralWaveModel", "InfiniteDepth", "HorizontalField", "Clock", "Simulation", "WENO", "set_", "time_step_", "run_",
    "interior", "PseudomomentumVelocities", "SpatialAveraging", "automatic_q_grid", "update_coupling_",
    "compute_pseudomomentum_cell_averages", "compute_pseudomomentum_cell_integrals", "pseudomomentum_fields",
    "compute_pseudomomentum_tendency_cell_averages", "pseudomomentum_tendency_fields", "cwcm_momentum_tendency_fields",
    "m0", "root_mean_square_wavenumber", "mean_direction", "significant_wave_height",
    "total_action", "total_deep_water_energy",
    "SymmetricQuadruplet", "DiscreteInteractionApproximation", "QuadrupletTransferInteraction", "ListedQuadruplets",
    "BulkWindDrag", "LinearStormTrack", "HollandHurricaneWind", "PressureCorrelationInput",
    "MeanSpectrumWhitecapping", "LocalSaturationDissipation", "WhitecappingDissipation", "PrecomputedSources",
    "wind_state",
    "SourceNetwork", "SurrogateConfig", "training_pairs", "attach_surrogate", "surrogate_stats",
    "train_surrogate", "Ensemble", "JLD2Writer", "IterationInterval", "TimeInterval", "fetch_output",
]
