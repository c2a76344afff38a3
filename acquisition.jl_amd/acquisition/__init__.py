"""MI355X FM-DBZP acquisition engine behind Acquisition.jl's public API.

Names mirror the reference module `Acquisition` (src/Acquisition.jl:12-29);
Julia's `acquire!` is spelled `acquire_` (a trailing underscore marks the
in-place call, as `!` does in Julia). Frequencies are plain Hz floats.
"""
from .estimators import AbstractNoiseEstimator, GlobalMeanNoiseEstimator, OppositeRowNoiseEstimator
from .plan import AcquisitionPlan, ArgumentError, BATCH_FFT_THRESHOLD, plan_acquire
from .cfar import cfar_threshold
from .acquire import (AcquisitionResults, Packed2BitSignal, acquire, acquire_, acquire_batch, get_num_cells,
                      is_detected)
from .systems import GNSSSystem, GPSL1CA, GPSL5I, E1Like, L5Like
from ._native import EngineError
from .recommend import (AbstractSDRClockPlan, AD9361ClockPlan, GpuCostModel, SamplingFreqRecommendation,
                        format_recommendations, is_valid_sample_rate, recommend_sampling_freqs, sample_rate_range)

__all__ = [
    "acquire", "acquire_", "acquire_batch", "plan_acquire", "AcquisitionPlan", "AcquisitionResults",
    "AbstractNoiseEstimator", "OppositeRowNoiseEstimator", "GlobalMeanNoiseEstimator",
    "cfar_threshold", "get_num_cells", "is_detected", "Packed2BitSignal", "ArgumentError", "EngineError",
    "GNSSSystem", "GPSL1CA", "GPSL5I", "E1Like", "L5Like", "BATCH_FFT_THRESHOLD",
    "recommend_sampling_freqs", "SamplingFreqRecommendation", "format_recommendations", "GpuCostModel",
    "AbstractSDRClockPlan", "AD9361ClockPlan", "is_valid_sample_rate", "sample_rate_range",
]
