"""Phasor convention helpers (src/beat_engine/julia_local/src/BeatEnginePhasor.jl:1-24).

exp(-i omega t) is the default; the signed wavenumber, the Neumann scale and the
Burton-Miller coupling all carry the propagation sign s = +1 / -1.
"""
NEGATIVE_TIME_PHASOR = "exp(-i omega t)"
POSITIVE_TIME_PHASOR = "exp(+i omega t)"


def propagation_sign(convention: str = NEGATIVE_TIME_PHASOR) -> int:
    if convention not in (NEGATIVE_TIME_PHASOR, POSITIVE_TIME_PHASOR):
        raise ValueError(f"Unsupported phasor convention: {convention!r}")
    return 1 if convention == NEGATIVE_TIME_PHASOR else -1


def outgoing_wavenumber(k: float, convention: str = NEGATIVE_TIME_PHASOR) -> float:
    return propagation_sign(convention) * abs(k)


def neumann_scale(density: float, omega: float, convention: str = NEGATIVE_TIME_PHASOR) -> complex:
    return complex(0.0, propagation_sign(convention) * density * omega)


def time_derivative(omega: float, convention: str = NEGATIVE_TIME_PHASOR) -> complex:
    return complex(0.0, -propagation_sign(convention) * omega)


def burton_miller_coupling(k: float, convention: str = NEGATIVE_TIME_PHASOR) -> complex:
    return complex(0.0, propagation_sign(convention)) / abs(k)
