"""Shape-factor models (the reference's public functions, host NumPy).

These are the callables ``SimulationGenerator(shape_factor_model=...)`` accepts
and stores (diffsims/utils/shape_factor_models.py).  Passing one of them
selects the matching HIP implementation (identified by function identity);
any other callable is evaluated on the host on the candidate reflections the
GPU selects.
"""

import numpy as np

__all__ = ["atanc", "binary", "linear", "lorentzian", "lorentzian_precession", "sin2c", "sinc"]


def binary(excitation_error, max_excitation_error):
    """Unit factor for every reflection (shape_factor_models.py:33-49)."""
    return 1


def linear(excitation_error, max_excitation_error):
    """max(1 - |s|/w, 0) (shape_factor_models.py:52-73)."""
    sf = 1 - np.abs(excitation_error) / max_excitation_error
    if isinstance(sf, np.ndarray):
        return np.maximum(sf, 0.0)
    return max(sf, 0.0)


def sinc(excitation_error, max_excitation_error, minima_number=5):
    """|sin(x)/x|, x = pi n s / w, and 0 where x = 0 (shape_factor_models.py:76-101)."""
    s = np.asarray(excitation_error, dtype=np.float64)
    fac = np.pi * minima_number / max_excitation_error
    num = np.sin(fac * s)
    den = fac * s
    out = np.zeros_like(num)
    np.divide(num, den, out=out, where=den != 0)
    return np.abs(out)


def sin2c(excitation_error, max_excitation_error, minima_number=5):
    """sinc^2 (shape_factor_models.py:104-123)."""
    return sinc(excitation_error, max_excitation_error, minima_number) ** 2


def atanc(excitation_error, max_excitation_error, minima_number=5):
    """atan(x)/x, x = pi n s / |w|, 1 at x = 0 (shape_factor_models.py:126-151)."""
    s = np.asarray(excitation_error, dtype=np.float64)
    fac = np.pi * minima_number / np.abs(max_excitation_error)
    with np.errstate(invalid="ignore", divide="ignore"):
        r = np.arctan(fac * s) / (fac * s)
    return np.nan_to_num(r, nan=1.0)


def lorentzian(excitation_error, max_excitation_error):
    """1 / (1 + (pi s / w)^2) (shape_factor_models.py:154-180)."""
    t = np.pi / max_excitation_error * np.asarray(excitation_error, dtype=np.float64)
    return 1.0 / (1.0 + t * t)


def lorentzian_precession(excitation_error, max_excitation_error, r_spot, precession_angle):
    """Precession-averaged Lorentzian (Palatinus et al. 2019, eq. 10;
    shape_factor_models.py:183-219); precession_angle in radians."""
    s = np.asarray(excitation_error, dtype=np.float64)
    sigma = np.pi / max_excitation_error
    u = sigma ** 2 * (np.asarray(r_spot) ** 2 * precession_angle ** 2 - s ** 2) + 1
    z = np.sqrt(u ** 2 + 4 * sigma ** 2 * s ** 2)
    return (sigma / np.pi) * np.sqrt(2 * (u + z) / z ** 2)
