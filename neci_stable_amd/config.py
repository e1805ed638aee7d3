"""Run specification: the CALC/SYSTEM values one engine needs, derived on the host.

Defaults follow SetCalcDefaults (Calc.F90:110-350): MEMORYFACPART 10,
MEMORYFACSPAWN 3, MEMORYFACHASH 0.7, ADDTOINITIATOR 3, REALSPAWNCUTOFF 0.95,
SETOCCUPIEDTHRESH 1, AvMCExcits 1.
"""
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from .systems import (HubbardChain, nwords, psingles_from_reference, random_hash2)

PART_UNIF_UNIF = 0
PART_UNIF_FAST = 1
PART_UNIF_FULL = 2
PART_FULL_FULL = 3
HOLES_FAST_FAST = 0
HOLES_FULL_FULL = 1


@dataclass
class Spec:
    system: object
    n_replicas: int = 1
    init_walkers: float = 1e4
    seed: int = 1234
    trunc_initiator: bool = True
    init_walk_no: float = 3.0
    all_real_coeff: bool = False
    real_spawn_cutoff: float = 0.95
    t_real_spawn_cutoff: bool = True
    occupied_thresh: float = 1.0
    av_mc_excits: float = 1.0
    memory_fac_part: float = 10.0
    memory_fac_spawn: float = 3.0
    memory_fac_hash: float = 0.7
    max_walkers: Optional[int] = None
    max_spawned: Optional[int] = None
    n_hash_buckets: Optional[int] = None
    singles: int = 1
    particles: int = PART_UNIF_FAST
    holes: int = HOLES_FAST_FAST  # PCHB hole selection: FAST-FAST or FULL-FULL (fully weighted)
    excitgen: str = "pchb"   # molecular: "pchb" (PCHB doubles + heat-bath singles) or "uniform" (gen_rand_excit)
    psingles: Optional[float] = None
    pparallel: float = 0.5
    ref_orbs: Optional[list] = None
    random_hash: Optional[np.ndarray] = None
    flag_layout: int = 0
    spawn_fp32: bool = False  # spawned amplitudes stored in FP32 (sums stay FP64)
    model: int = field(init=False)
    nel: int = field(init=False)
    nbasis: int = field(init=False)
    W: int = field(init=False)
    S: int = field(init=False)

    def __post_init__(self):
        s = self.system
        self.nel = s.nel
        self.nbasis = s.nbasis
        self.W = nwords(self.nbasis)
        self.S = int(self.n_replicas)
        self.model = 0 if isinstance(s, HubbardChain) else 1
        if self.max_walkers is None:
            self.max_walkers = int(self.memory_fac_part * self.init_walkers) + 1024
        if self.max_spawned is None:
            self.max_spawned = int(self.memory_fac_spawn * self.init_walkers * self.S) + 1024
        if self.n_hash_buckets is None:
            self.n_hash_buckets = max(1, int(round(self.memory_fac_hash * self.init_walkers)))
        if self.ref_orbs is None:
            self.ref_orbs = s.neel_state() if self.model == 0 else s.aufbau_reference()
        self.ref_orbs = sorted(int(o) for o in self.ref_orbs)
        if self.model == 0:
            self.psingles = 1.0
        elif self.psingles is None:
            self.psingles = psingles_from_reference(self.ref_orbs, self.nbasis)
        if self.random_hash is None:
            self.random_hash = random_hash2(self.nbasis, self.seed)
        self.random_hash = np.ascontiguousarray(self.random_hash, dtype=np.int32)

    @property
    def params(self):
        return dict(trunc_initiator=self.trunc_initiator, init_walk_no=self.init_walk_no,
                    all_real_coeff=self.all_real_coeff, real_spawn_cutoff=self.real_spawn_cutoff,
                    t_real_spawn_cutoff=self.t_real_spawn_cutoff, occupied_thresh=self.occupied_thresh,
                    av_mc_excits=self.av_mc_excits)
