"""MI355X-native gauge+Higgs HMC molecular dynamics (drop-in for the gh_hmc /
hmc_higgs path of hlammiv/gauge_higgs). The compute lives in libghx.so (HIP,
gfx950); this package is the host-side mirror of the reference interface."""
from .hmc import (COLD, FP32MD, FP64, HOT, LEAPFROG, OMELYAN, GaugeHiggsHMC, GaugeHMC, GhxError, Lattices, autocorr_function, casimir_channels,
                  rng_gauss, series_analysis)
from .rep import rep_query
from .replicas import ReplicaSet, kappa_scan

__all__ = ["ReplicaSet", "kappa_scan", "GaugeHMC", "GaugeHiggsHMC", "Lattices", "GhxError", "rep_query", "casimir_channels", "series_analysis", "autocorr_function", "rng_gauss", "FP64", "FP32MD", "LEAPFROG", "OMELYAN", "HOT", "COLD"]
