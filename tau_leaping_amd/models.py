"""Reaction networks as channel-major (CSC) stoichiometry, and the synthetic
model families used by the benchmarks.

``Network.to_csc`` follows ReadXMLToMatrixStruct
(tau_leaping/CSCHostMatrixData.cpp:19-107): each channel's entries are sorted
by species index, reactants carry -stoichiometry and products +stoichiometry,
and a species that is both reactant and product gets two entries (reactant
first).  Species index = order of declaration (StochKitReader.cpp:106-121).

The generators reproduce the structure of toy_system_gen
(toy_system_gen/main.cpp:654-756: random networks, cyclic chain ``-w``,
colloidal aggregation ``-s``) with explicit seeds and the rate / population
recipes of the benchmark configurations.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Sequence

import numpy as np


@dataclass
class Reaction:
    reactants: Dict[int, int]
    products: Dict[int, int]
    rate: float


@dataclass
class Network:
    n_species: int
    reactions: List[Reaction]
    x0: np.ndarray
    species_names: List[str] = field(default_factory=list)
    name: str = "network"

    @property
    def n_channels(self) -> int:
        return len(self.reactions)

    def to_csc(self):
        """Return (col_ptr, row_ind, stoich, rates, x0) arrays."""
        col_ptr = [0]
        row_ind: List[int] = []
        stoich: List[int] = []
        for rx in self.reactions:
            ind = sorted(list(rx.reactants.keys()) + list(rx.products.keys()))
            found = set()
            for s in ind:
                if s in rx.reactants and s not in found:
                    stoich.append(-int(rx.reactants[s]))
                    found.add(s)
                else:
                    stoich.append(int(rx.products[s]))
                row_ind.append(s)
            col_ptr.append(len(row_ind))
        rates = np.array([rx.rate for rx in self.reactions], np.float64)
        return (np.array(col_ptr, np.int32), np.array(row_ind, np.int32), np.array(stoich, np.int8),
                rates, np.ascontiguousarray(self.x0, np.int32))

    def names(self) -> List[str]:
        return self.species_names or [f"S{i + 1}" for i in range(self.n_species)]

    def to_stochkit_xml(self) -> str:
        """StochKit 2.0 text (the subset StochKitReader.cpp:35-179 reads)."""
        names = self.names()
        out = ["<Model>", f"  <Description>{self.name}</Description>",
               f"  <NumberOfReactions>{self.n_channels}</NumberOfReactions>",
               f"  <NumberOfSpecies>{self.n_species}</NumberOfSpecies>", "  <ParametersList>"]
        for j, rx in enumerate(self.reactions):
            out += ["    <Parameter>", f"      <Id>c{j + 1}</Id>",
                    f"      <Expression>{rx.rate!r}</Expression>", "    </Parameter>"]
        out += ["  </ParametersList>", "  <ReactionsList>"]
        for j, rx in enumerate(self.reactions):
            out += ["    <Reaction>", f"      <Id>R{j + 1}</Id>", "      <Type>mass-action</Type>",
                    f"      <Rate>c{j + 1}</Rate>", "      <Reactants>"]
            for s, v in rx.reactants.items():
                out.append(f'        <SpeciesReference id="{names[s]}" stoichiometry="{v}"/>')
            out += ["      </Reactants>", "      <Products>"]
            for s, v in rx.products.items():
                out.append(f'        <SpeciesReference id="{names[s]}" stoichiometry="{v}"/>')
            out += ["      </Products>", "    </Reaction>"]
        out += ["  </ReactionsList>", "  <SpeciesList>"]
        for i in range(self.n_species):
            out += ["    <Species>", f"      <Id>{names[i]}</Id>",
                    f"      <InitialPopulation>{int(self.x0[i])}</InitialPopulation>", "    </Species>"]
        out += ["  </SpeciesList>", "</Model>", ""]
        return "\n".join(out)


def dimer_decay(s1: int = 100000) -> Network:
    """examples/dimer_decay.xml (c = 1.0, 0.002, 0.5, 0.04) with S1 overridable."""
    rx = [Reaction({0: 1}, {}, 1.0), Reaction({0: 2}, {1: 1}, 0.002),
          Reaction({1: 1}, {0: 2}, 0.5), Reaction({1: 1}, {2: 1}, 0.04)]
    return Network(3, rx, np.array([s1, 0, 0], np.int32), ["S1", "S2", "S3"], "dimer_decay")


def gene_expression(genes: int = 50) -> Network:
    """A network with catalysts (species that are reactant and product of one
    channel): transcription G -> G + M, translation M -> M + P, decays,
    dimerisation 2P <-> D, repression D + G -> D + Gi with release Gi -> G,
    and 2P -> P (net coefficient -1 of P)."""
    rx = [Reaction({0: 1}, {0: 1, 1: 1}, 5.0), Reaction({1: 1}, {1: 1, 2: 1}, 2.0),
          Reaction({1: 1}, {}, 1.0), Reaction({2: 1}, {}, 0.05),
          Reaction({2: 2}, {3: 1}, 1e-5), Reaction({3: 1}, {2: 2}, 0.1),
          Reaction({0: 1, 3: 1}, {3: 1, 4: 1}, 1e-5), Reaction({4: 1}, {0: 1}, 0.5),
          Reaction({2: 2}, {2: 1}, 1e-6)]
    return Network(5, rx, np.array([genes, 250, 5000, 1000, 0], np.int32), ["G", "M", "P", "D", "Gi"],
                   "gene_expression")


def _log_uniform(rng, lo, hi, size=None):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size))


def toy_network(n_species: int = 128, n_channels: int = 256, seed: int = 42, rate_lo: float = 1e-3,
                rate_hi: float = 1.0, pop_lo: int = 0, pop_hi: int = 10000) -> Network:
    """Random network: per channel 1-2 distinct reactants (coefficient 1) drawn
    uniformly, 1-2 distinct products different from the reactants
    (coefficient 1); rates log-uniform; populations uniform integers."""
    rng = np.random.default_rng(seed)
    rx = []
    for _ in range(n_channels):
        nr = int(rng.integers(1, 3))
        reac = rng.choice(n_species, size=nr, replace=False)
        npr = int(rng.integers(1, 3))
        pool = np.setdiff1d(np.arange(n_species), reac)
        prod = rng.choice(pool, size=npr, replace=False)
        rx.append(Reaction({int(s): 1 for s in reac}, {int(s): 1 for s in prod},
                           float(_log_uniform(rng, rate_lo, rate_hi))))
    x0 = rng.integers(pop_lo, pop_hi + 1, size=n_species).astype(np.int32)
    return Network(n_species, rx, x0, name=f"toy_{n_species}_{n_channels}")


def cyclic_chain(n_species: int = 1024, seed: int = 43, rate: float = 1.0, pop_lo: int = 1000,
                 pop_hi: int = 2000) -> Network:
    """Weakly coupled cyclic chain S_i -> S_(i+1) mod N (toy_system_gen -w,
    main.cpp:694-710)."""
    rng = np.random.default_rng(seed)
    rx = [Reaction({i: 1}, {(i + 1) % n_species: 1}, float(rate)) for i in range(n_species)]
    x0 = rng.integers(pop_lo, pop_hi + 1, size=n_species).astype(np.int32)
    return Network(n_species, rx, x0, name=f"cyclic_{n_species}")


def colloidal(n_species: int = 250, seed: int = 44, rate_lo: float = 1e-4, rate_hi: float = 1e-1,
              pop_lo: int = 1000, pop_hi: int = 2000) -> Network:
    """Strongly coupled colloidal aggregation (toy_system_gen -s,
    main.cpp:712-756): S_n + S_m -> S_(n+m) for n <= m, n+m <= N, then
    fragmentation S_p -> S_q + S_(p-q) for q <= p/2.  M = N^2/2."""
    rng = np.random.default_rng(seed)
    rx = []
    N = n_species
    for n in range(1, N // 2 + 1):
        rx.append(Reaction({n - 1: 2}, {2 * n - 1: 1}, 0.0))
        for m in range(n + 1, N - n + 1):
            rx.append(Reaction({n - 1: 1, m - 1: 1}, {m + n - 1: 1}, 0.0))
    for p in range(1, N + 1):
        for q in range(1, p // 2 + 1):
            if q == p - q:
                rx.append(Reaction({p - 1: 1}, {q - 1: 2}, 0.0))
            else:
                rx.append(Reaction({p - 1: 1}, {q - 1: 1, p - q - 1: 1}, 0.0))
    rates = _log_uniform(rng, rate_lo, rate_hi, len(rx))
    for r, c in zip(rx, rates):
        r.rate = float(c)
    x0 = rng.integers(pop_lo, pop_hi + 1, size=N).astype(np.int32)
    return Network(N, rx, x0, name=f"colloid_{N}")


def hetero_batch(n_models: int = 64, base_seed: int = 1000, n_lo: int = 16, n_hi: int = 1024,
                 r_lo: int = 256, r_hi: int = 8192):
    """Heterogeneous batch: species log-uniform in [n_lo, n_hi], channels
    uniform in [2N, 4N], realizations log-uniform in [r_lo, r_hi]; the toy
    recipe per model with seed base_seed + m.  Returns (networks, R list)."""
    nets, reals = [], []
    for m in range(n_models):
        rng = np.random.default_rng(base_seed + m)
        n = int(round(float(_log_uniform(rng, n_lo, n_hi))))
        ch = int(rng.integers(2 * n, 4 * n + 1))
        r = int(round(float(_log_uniform(rng, r_lo, r_hi))))
        nets.append(toy_network(n, ch, seed=base_seed + m))
        reals.append(r)
    return nets, reals


def from_csc(n_species: int, col_ptr: Sequence[int], row_ind: Sequence[int], stoich: Sequence[int],
             rates: Sequence[float], x0: Sequence[int]) -> Network:
    """Inverse of Network.to_csc (catalyst double entries are merged back)."""
    rx = []
    for j in range(len(col_ptr) - 1):
        reac: Dict[int, int] = {}
        prod: Dict[int, int] = {}
        for e in range(col_ptr[j], col_ptr[j + 1]):
            s, v = int(row_ind[e]), int(stoich[e])
            if v < 0:
                reac[s] = -v
            elif v > 0:
                prod[s] = v
        rx.append(Reaction(reac, prod, float(rates[j])))
    return Network(int(n_species), rx, np.asarray(x0, np.int32))
