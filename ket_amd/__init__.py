"""ket_amd — MI355X-native dense statevector backend for ket.

The public surface mirrors ket's (src/index.ts:1-13) for the statevector
path: Circuit, Distribution, the gate matrices and the algorithms built on
the path.  All simulation runs in libketsv.so (HIP, gfx950) through the C
ABI of include/ketsv.h.
"""
from . import gates
from ._native import RangeError, State, device_count
from .algorithms import (AnsatzFn, efficientSU2, energy_and_gradient, energy_and_gradient_batch, gradient, iqft,
                         maxCutHamiltonian, minimize, qaoa, qft, realAmplitudes, vqe, vqe_batch)
from .circuit import Circuit, Distribution, configure, make_prng, prng_draws
from .noise import DEVICES

__all__ = ["Circuit", "Distribution", "RangeError", "State", "gates", "qft", "iqft", "qaoa", "maxCutHamiltonian",
           "realAmplitudes", "efficientSU2", "vqe", "gradient", "energy_and_gradient", "energy_and_gradient_batch", "vqe_batch", "minimize", "AnsatzFn",
           "configure", "make_prng", "prng_draws", "DEVICES", "device_count"]
