"""ORACLE (test infrastructure only): the host facts a CPU baseline states
(CPU model, logical CPU count, the core it was pinned to) and single-core
pinning for the duration of a timing, the in-process equivalent of
`taskset -c <core>` (sched_setaffinity on this process)."""
from __future__ import annotations

import contextlib
import os
import platform


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def nproc() -> int:
    """Logical CPUs this process may run on (what `nproc` prints)."""
    return len(os.sched_getaffinity(0))


@contextlib.contextmanager
def pinned(core: int | None = None):
    """Pin the calling process to one logical CPU (the last one it may use, or
    `core`) and restore the previous mask on exit.  Yields the host facts a
    baseline reports: CPU model, nproc before pinning, the pinned core."""
    old = os.sched_getaffinity(0)
    c = max(old) if core is None else core
    facts = {"cpu_model": cpu_model(), "nproc": len(old), "pinned_core": c,
             "pinning": f"sched_setaffinity to CPU {c} (as taskset -c {c})"}
    os.sched_setaffinity(0, {c})
    try:
        yield facts
    finally:
        os.sched_setaffinity(0, old)
