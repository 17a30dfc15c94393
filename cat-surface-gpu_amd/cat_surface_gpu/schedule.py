"""Launch bounds and deferred checks for a registration schedule that runs
either eagerly or inside one HIP graph.

Two things in the CAT_SurfWarp stages depend on data the host only sees after a
synchronisation: the heat-kernel iteration counts (from each mesh's mean edge
length, surface_stencil.py:349-370) and the checks that a step's result is
valid (stencil index capacity and reach, SOR barrier time-outs and XCD
placement, sphere normalisation, rotation coverage).  The kernels read the
counts from device memory (csg_heat_parameters and the ``_dev`` smoothing
entries), so the host only has to choose how many launches to issue:

* eagerly, ``launch_bound`` reads the counts (one synchronisation) and issues
  exactly that many, and ``defer`` checks a flag at once;
* while a schedule is recorded (``recording``), the same happens and every
  bound is kept in call order;
* while it is captured (``capturing``), the bounds come from that record with
  a margin, every launch past a row's count copies the row through, and each
  bound and flag becomes a device check that the caller reads after the
  graph's replay (``Capture.verdict``): any failure means the replay's output
  must not be used.
"""

from __future__ import annotations

import math
from contextlib import contextmanager

import torch

from . import _lib

_STACK: list = []


class Recording:
    """Bounds of an eager run, in call order."""

    def __init__(self) -> None:
        self.bounds: list[int] = []


class Capture:
    """Bounds for a captured run (from a Recording, with a margin) and the
    device flags it leaves to check."""

    def __init__(self, recorded: list[int], margin: float = 1.25, slack: int = 8) -> None:
        self.bounds = [int(math.ceil(b * margin)) + slack for b in recorded]
        self.position = 0
        self.flags: list[torch.Tensor] = []
        self.labels: list[str] = []

    def next_bound(self) -> int:
        if self.position >= len(self.bounds):
            raise RuntimeError("the captured schedule issues more bounded loops than the "
                               "recorded one")
        b = self.bounds[self.position]
        self.position += 1
        return b

    def verdict(self) -> list[str]:
        """Labels of the checks that failed in the last replay (one host read)."""

        if not self.flags:
            return []
        host = torch.cat([f.reshape(-1).to(torch.int32) for f in self.flags]).cpu().tolist()
        sizes = [int(f.numel()) for f in self.flags]
        failed, pos = [], 0
        for label, size in zip(self.labels, sizes):
            if any(host[pos:pos + size]):
                failed.append(label)
            pos += size
        return failed


def current():
    return _STACK[-1] if _STACK else None


def capturing() -> bool:
    return isinstance(current(), Capture)


@contextmanager
def recording():
    rec = Recording()
    _STACK.append(rec)
    try:
        yield rec
    finally:
        _STACK.pop()


@contextmanager
def capture(recorded: list[int], margin: float = 1.25):
    cap = Capture(recorded, margin)
    _STACK.append(cap)
    try:
        yield cap
    finally:
        _STACK.pop()


def _exceeds(counts: list[torch.Tensor], bound: int) -> torch.Tensor:
    """Device int32 [1]: 1 when some count exceeds ``bound``."""

    flag = torch.zeros(1, dtype=torch.int32, device=counts[0].device)
    for c in counts:
        c = c.reshape(-1).to(torch.int32).contiguous()
        _lib.check(_lib.lib().csg_flag_above(_lib.ptr(c), int(c.numel()), int(bound),
                                             _lib.ptr(flag), _lib.stream_handle(c.device)),
                   "csg_flag_above")
    return flag


def launch_bound(counts: list[torch.Tensor], label: str = "heat iterations") -> int:
    """How many iterations to launch for rows with these device counts."""

    state = current()
    if isinstance(state, Capture):
        bound = state.next_bound()
        state.flags.append(_exceeds(counts, bound))
        state.labels.append(f"{label} above {bound}")
        return bound
    flat = torch.cat([c.reshape(-1).to(torch.int64) for c in counts])
    bound = int(flat.max().item()) if flat.numel() else 0
    if isinstance(state, Recording):
        state.bounds.append(bound)
    return bound


def defer(flag: torch.Tensor, label: str, error=RuntimeError) -> None:
    """A device flag that must be zero: checked now when eager, after the
    replay when captured."""

    state = current()
    if isinstance(state, Capture):
        state.flags.append(flag)
        state.labels.append(label)
        return
    if bool((flag.reshape(-1) != 0).any().item()):
        raise error(label)
