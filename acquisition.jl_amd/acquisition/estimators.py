"""Noise-power estimator selectors (src/est_signal_noise_power.jl:30-100)."""


class AbstractNoiseEstimator:
    """Marks which noise-power estimator `acquire_` uses for the CFAR test statistic."""

    def __eq__(self, other):
        return type(self) is type(other)

    def __hash__(self):
        return hash(type(self))

    def __repr__(self):
        return f"{type(self).__name__}()"


class OppositeRowNoiseEstimator(AbstractNoiseEstimator):
    """Default: mean of the Doppler row half the grid away from the peak."""


class GlobalMeanNoiseEstimator(AbstractNoiseEstimator):
    """Mean of every cell except +-samples_per_chip columns around the peak."""
