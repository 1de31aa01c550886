"""Autotuner of the particle phase: picks the tile shape (cells, and so
threads, per workgroup: 4x4x4 = 256 or 8x4x4 = 512) and the sort schedule
(rebuild the tile segments every N phases, or only on load) by timing each
candidate on the caller's own workload -- a batch of independent instances
(tune_batch) or one engine (tune_engine) -- and, for an FP32 fixed-point
engine, the particles each thread pushes per batch of the phase
(tune_engine's ``ppt_choices``: one, or two with buckets twice as deep, which
costs the third workgroup per compute unit).  Batches run one particle per
thread.

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
    ppt: int = 1  # particles per thread and batch


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
                        "sort_every": t.sort_every, "particles_per_thread": t.ppt,
                        "ms_per_phase": t.ms_per_phase}
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
    ppt: int = 1

    def as_dict(self) -> dict:
        d = TuneResult(self.tile, self.sort_every, self.ms_per_phase, self.trials).as_dict()
        d["particles_per_thread"] = self.ppt
        return d


def tune_engine(make_runner: Callable[..., object],
                tiles: Sequence[Tuple[int, int, int]] = TILES,
                sort_choices: Sequence[int] = SORT_CHOICES, phases: int = 10,
                warmup: int = 1, stream: Optional[Stream] = None,
                ppt_choices: Optional[Sequence[int]] = None) -> EngineTuneResult:
    """Tile shape, sort interval and (with ``ppt_choices``) particles per
    thread of one engine's phase.

    ``make_runner(tile, sort_every)`` -- ``make_runner(tile, sort_every,
    ppt)`` when ``ppt_choices`` is given -- builds a fresh engine with those
    settings, loaded from the caller's initial condition, and returns an
    object with ``phase(stream)`` (one engine phase, the engine's own
    sorts included) and ``close()``.  Each candidate runs ``warmup`` untimed
    and ``phases`` timed phases on one stream in async mode, timed with HIP
    events; the fastest wins."""
    if not tiles or not sort_choices or (ppt_choices is not None and not ppt_choices):
        raise ValueError("tune_engine needs at least one choice of each setting")
    if phases < 1 or warmup < 0:
        raise ValueError("tune_engine needs phases >= 1 and warmup >= 0")
    stream = stream or Stream()
    trials: List[Trial] = []
    candidates = [(tuple(tile), int(se), ppt)
                  for ppt in (ppt_choices if ppt_choices is not None else (None,))
                  for tile in tiles for se in sort_choices]
    for tile, se, ppt in candidates:
        run = make_runner(tile, se) if ppt is None else make_runner(tile, se, int(ppt))
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
        trials.append(Trial(tuple(tile), int(se), a.elapsed_ms(b) / phases,
                            1 if ppt is None else int(ppt)))
    best = min(trials, key=lambda t: t.ms_per_phase)
    return EngineTuneResult(best.tile, best.sort_every, best.ms_per_phase, trials, best.ppt)
