"""Result containers with the fields of pycwb/modules/likelihoodWP/typing.py."""

from dataclasses import dataclass

import numpy as np


@dataclass
class SkyStatistics:
    """Statistics at the selected sky direction (typing.py:5-69)."""
    Gn: np.float32
    Ec: np.float32
    Dc: np.float32
    Rc: np.float32
    Eh: np.float32
    Es: np.float32
    Np: np.float32
    Em: np.float32
    Lm: np.float32
    norm: np.float32
    cc: np.float32
    rho: np.float32
    xrho: np.float32
    Lo: np.float32
    Eo: np.float32
    N_pix_effective: np.float32
    energy_array_plus: np.ndarray
    energy_array_cross: np.ndarray
    v00: np.ndarray
    v90: np.ndarray
    pd: np.ndarray
    pD: np.ndarray
    ps: np.ndarray
    pS: np.ndarray
    pixel_mask: np.ndarray
    gaussian_noise_correction: np.ndarray
    noise_amplitude_00: np.ndarray
    noise_amplitude_90: np.ndarray
    coherent_energy: np.ndarray
    p00_POL: np.ndarray
    p90_POL: np.ndarray
    r00_POL: np.ndarray
    r90_POL: np.ndarray
    S_snr: np.ndarray
    f: np.ndarray
    F: np.ndarray


@dataclass
class SkyMapStatistics:
    """Sky maps of one cluster (typing.py:72-98)."""
    l_max: int
    nAntennaPrior: np.ndarray
    nAlignment: np.ndarray
    nLikelihood: np.ndarray
    nNullEnergy: np.ndarray
    nCorrEnergy: np.ndarray
    nCorrelation: np.ndarray
    nSkyStat: np.ndarray
    nDisbalance: np.ndarray
    nNetIndex: np.ndarray
    nEllipticity: np.ndarray
    nPolarisation: np.ndarray
    sky_stat_max: float = 0
    nProbability: np.ndarray = None
    stage_timings: dict = None

    @classmethod
    def from_tuple(cls, t):
        return cls(*t)
