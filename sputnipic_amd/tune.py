"""Autotuner of the particle phase: picks the tile shape (cells, and so
threads, per workgroup: 4x4x4 = 256 or 8x4x4 = 512) and the sort schedule
(rebuild the tile segments every N phases, or only on load) by timing each
candidate on the caller's own workload -- a batch of independent instances
(tune_batch) or one engine (tune_engine).  Particles per thread are one per
batch: the LDS budget of the tile (field stage, FP64 moment block, per-cell
buckets) holds exactly one batch of buckets at three workgroups per compute
unit, so the workgroup shape is the tuned launch dimension.

Each tile shape needs its own engines, so the caller supplies a builder:
``make_instances(tile)`` returns fresh instances (loaded from the same
initial condition every time) with ``member()`` (the Batch member tuple),
``zero(stream)`` (zero the instance's moments) and ``close()``.  Every
candidate starts from freshly sorted instances (every member engine is
sorted first, so that no candidate inherits the order another one let
decay) and runs ``warmup`` untimed and ``phases`` timed cycles (zero unless
the batch writes the moments, batch phase) on one stream, timed with HIP
events; the schedule's sorts fall inside that span when ``warmup +
phases`` exceeds the sort interval.  The tuner advances only its own
instances; the caller builds the production instances afterwards with the
winning settings.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence, Tuple

from .core import Batch, Event, Stream, check_domain, get_async, set_async

TILES = ((4, 4, 4), (8, 4, 4))
SORT_CHOICES = (0, 10, 5)


@dataclass
class Trial:
    tile: Tuple[int, int, int]
    sort_every: int
    ms_per_phase: float


@dataclass
class TuneResult:
    tile: Tuple[int, int, int]
    sort_every: int
    ms_per_phase: float
    trials: List[Trial] = field(default_factory=list)

    def as_dict(self) -> dict:
        return {
            "tile": "x".join(str(t) for t in self.tile),
            "sort_every": self.sort_every,
            "ms_per_phase": self.ms_per_phase,
            "trials": [{"tile": "x".join(str(t) for t in t.tile),
                        "sort_every": t.sort_every, "ms_per_phase": t.ms_per_phase}
                       for t in self.trials],
        }


def time_batch(insts, batch: Batch, stream: Stream, phases: int, warmup: int) -> float:
    """Mean milliseconds per cycle (zero the moments unless the batch writes
    them, batch phase), with the calls in async mode so that no host wait
    falls inside the timed span."""
    def cycle():
        if not batch.set_moments:
            for x in insts:
                x.zero(stream)
        batch.phase(stream)

    was_async = get_async()
    set_async(True)
    try:
        for _ in range(warmup):
            cycle()
        a, b = Event(), Event()
        a.record(stream)
        for _ in range(phases):
            cycle()
        b.record(stream)
        stream.synchronize()
        check_domain(stream)
    finally:
        set_async(was_async)
    return a.elapsed_ms(b) / phases


def tune_batch(make_instances: Callable[[Tuple[int, int, int]], Sequence],
               tiles: Sequence[Tuple[int, int, int]] = TILES,
               sort_choices: Sequence[int] = SORT_CHOICES, phases: int = 10,
               warmup: int = 2, stream: Optional[Stream] = None,
               set_moments: bool = False) -> TuneResult:
    if not tiles or not sort_choices:
        raise ValueError("tune_batch needs at least one tile shape and one sort choice")
    if phases < 1 or warmup < 0:
        raise ValueError("tune_batch needs phases >= 1 and warmup >= 0")
    stream = stream or Stream()
    trials: List[Trial] = []
    for tile in tiles:
        insts = make_instances(tuple(tile))
        try:
            for se in sort_choices:
                for x in insts:
                    x.member()[0].sort(stream)
                batch = Batch([x.member() for x in insts], sort_every=se,
                              set_moments=set_moments)
                try:
                    ms = time_batch(insts, batch, stream, phases, warmup)
                finally:
                    batch.close()
                trials.append(Trial(tuple(tile), int(se), ms))
        finally:
            for x in insts:
                x.close()
    best = min(trials, key=lambda t: t.ms_per_phase)
    return TuneResult(best.tile, best.sort_every, best.ms_per_phase, trials)


@dataclass
class EngineTuneResult:
    tile: Tuple[int, int, int]
    sort_every: int
    ms_per_phase: float
    trials: List[Trial] = field(default_factory=list)

    def as_dict(self) -> dict:
        return TuneResult(self.tile, self.sort_every, self.ms_per_phase,
                          self.trials).as_dict()


def tune_engine(make_runner: Callable[[Tuple[int, int, int], int], object],
                tiles: Sequence[Tuple[int, int, int]] = TILES,
                sort_choices: Sequence[int] = SORT_CHOICES, phases: int = 10,
                warmup: int = 1, stream: Optional[Stream] = None) -> EngineTuneResult:
    """Tile shape and sort interval of one engine's phase.

    ``make_runner(tile, sort_every)`` builds a fresh engine with those
    settings, loaded from the caller's initial condition, and returns an
    object with ``phase(stream)`` (one engine phase, the engine's own
    sorts included) and ``close()``.  Each candidate runs ``warmup`` untimed
    and ``phases`` timed phases on one stream in async mode, timed with HIP
    events; the fastest wins."""
    if not tiles or not sort_choices:
        raise ValueError("tune_engine needs at least one tile shape and one sort choice")
    if phases < 1 or warmup < 0:
        raise ValueError("tune_engine needs phases >= 1 and warmup >= 0")
    stream = stream or Stream()
    trials: List[Trial] = []
    for tile in tiles:
        for se in sort_choices:
            run = make_runner(tuple(tile), int(se))
            was_async = get_async()
            set_async(True)
            try:
                for _ in range(warmup):
                    run.phase(stream)
                a, b = Event(), Event()
                a.record(stream)
                for _ in range(phases):
                    run.phase(stream)
                b.record(stream)
                stream.synchronize()
                check_domain(stream)
            finally:
                set_async(was_async)
                run.close()
            trials.append(Trial(tuple(tile), int(se), a.elapsed_ms(b) / phases))
    best = min(trials, key=lambda t: t.ms_per_phase)
    return EngineTuneResult(best.tile, best.sort_every, best.ms_per_phase, trials)
