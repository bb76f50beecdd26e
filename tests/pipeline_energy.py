"""Ground-state energy from a Trotter overlap series, as the reference's
examples pipeline computes it (`ph2 energy fft`, util/_ph2/analysis.py:
162-176): values[s] = <psi|U^(s+1)|psi> at spacing delta, the dominant
non-DC bin of |FFT(values)| at signed frequency f gives

    E = f * tau / (delta * normalization) + offset.

The reference's CI accepts the pipeline when -80 < E < -70 Ha for the
bundled Hamiltonian at delta 0.1 and 256 steps
(.github/workflows/CI.yml:44-57)."""

import math

import numpy as np

BAND = (-80.0, -70.0)


def fft_energy(values, delta, normalization, offset):
    v = np.asarray(values, dtype=np.complex128)
    amp = np.abs(np.fft.fft(v))
    amp[0] = 0.0                  # the off-resonant baseline
    f = np.fft.fftfreq(len(v))[int(np.argmax(amp))]
    return float(f / (normalization * delta / math.tau) + offset)
