"""Seeded synthetic workloads for the batched and served solve forms.

batched_instances: qubit models of the local-only driven family,
``driven_qubits(N; drive=2*Omega, detuning=1, emission=gamma,
dephasing=4*gamma_phi)`` (src/model_zoo.jl:214-236), with N drawn from
{20, 40, ..., 120}, Omega ~ U[0, 2], gamma_phi log-uniform on [1e-3, 1e-1],
gamma = 0.1, from numpy's PCG64 with a fixed seed.

served_requests: a Poisson arrival stream of mixed requests (60 % stationary
states, 40 % dynamics to t = 10 with 50 save points) on qubit models with
N ~ U{16..160}, Omega in [0, 2], gamma in [0.05, 0.2], gamma_phi in [0, 0.1].
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

BATCH_SEED = 20261203


@dataclass(frozen=True)
class Instance:
    N: int
    omega: float
    gamma: float
    gamma_phi: float

    def model(self):
        from .api import Models
        return Models.driven_qubits(self.N, drive=2 * self.omega, detuning=1.0, emission=self.gamma,
                                    dephasing=4 * self.gamma_phi, pumping=0.0)


def batched_instances(count=512, seed=BATCH_SEED, sizes=(20, 40, 60, 80, 100, 120)):
    """`count` seeded instances; N drawn uniformly from `sizes` (a single size
    gives a same-N batch with the same parameter distribution)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(count):
        N = int(rng.choice(list(sizes)))
        omega = float(rng.uniform(0.0, 2.0))
        gphi = float(np.exp(rng.uniform(np.log(1e-3), np.log(1e-1))))
        out.append(Instance(N, omega, 0.1, gphi))
    return out


@dataclass(frozen=True)
class Request:
    arrival: float      # seconds after the stream starts
    kind: str           # "steady" or "dynamics"
    instance: Instance


def served_requests(rate, count, seed=BATCH_SEED + 1, nmin=16, nmax=160):
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    out = []
    for _ in range(count):
        t += float(rng.exponential(1.0 / rate))
        kind = "steady" if rng.uniform() < 0.6 else "dynamics"
        N = int(rng.integers(nmin, nmax + 1))
        inst = Instance(N, float(rng.uniform(0.0, 2.0)), float(rng.uniform(0.05, 0.2)),
                        float(rng.uniform(0.0, 0.1)))
        out.append(Request(t, kind, inst))
    return out
