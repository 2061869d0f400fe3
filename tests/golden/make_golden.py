"""Writes the committed golden fixtures for the oracle and engine tests.

Everything here comes from numpy (pocketfft) and scipy, i.e. independent
library implementations of the FFT, erfc and adaptive quadrature; nothing is
taken from the reference (which ships no fixtures, no tests and no data
beyond an output-time list).  Run from the repository root:

    python tests/golden/make_golden.py
"""
import math
import os

import numpy as np
from scipy import integrate, special

HERE = os.path.dirname(os.path.abspath(__file__))


def fft_vectors():
    rng = np.random.default_rng(20240601)
    out = {}
    for n in (7, 8, 12, 30, 64):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out[f"c1d_{n}_in"] = x
        out[f"c1d_{n}_fwd"] = np.fft.fft(x)
    for M in (8, 12, 16):
        a = rng.standard_normal((M, M, M))
        out[f"r3d_{M}_in"] = a
        out[f"r3d_{M}_rfftn"] = np.fft.rfftn(a)
    np.savez_compressed(os.path.join(HERE, "fft_golden.npz"), **out)


def erfc_table():
    # Short-range factor erfc(u) + 2u/sqrt(pi) exp(-u^2) at the 1000 bin
    # centres u = 3/1000*(i+0.5), from scipy.special.erfc.
    u = 3.0 / 1000 * (np.arange(1000) + 0.5)
    t = special.erfc(u) + 2.0 * u / math.sqrt(math.pi) * np.exp(-u * u)
    np.savez_compressed(os.path.join(HERE, "erfc_table.npz"), u=u, table=t)


def drift_kick():
    # LCDM drift/kick integrals by scipy.integrate.quad (QUADPACK qags).
    om, ol, hub = 0.3, 0.7, 0.1
    a0, a1 = 0.02, 0.05

    def H(a):
        return hub * math.sqrt(om / a ** 3 + (1 - om - ol) / a ** 2 + ol)

    avals = np.exp(np.linspace(math.log(a0), math.log(a1), 11))[1:]
    drift = [integrate.quad(lambda a: 1 / (H(a) * a ** 3), a0, a, epsabs=0,
                            epsrel=1e-13, limit=200)[0] for a in avals]
    kick = [integrate.quad(lambda a: 1 / (H(a) * a ** 2), a0, a, epsabs=0,
                           epsrel=1e-13, limit=200)[0] for a in avals]
    np.savez_compressed(os.path.join(HERE, "drift_kick.npz"), a0=a0, a1=a1,
                        avals=avals, drift=np.array(drift),
                        kick=np.array(kick), omega0=om, omega_lambda=ol,
                        hubble=hub)


if __name__ == "__main__":
    fft_vectors()
    erfc_table()
    drift_kick()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))
