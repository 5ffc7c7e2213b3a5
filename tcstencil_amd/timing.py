"""Warm-up before a timed GPU region, shared by the benchmark drivers.

An idle MI355X raises its clocks over the first tens of milliseconds of load.
box2d2r 16384^2 FP16 sweeps (4x8-brick tiles) timed in chunks of 20 from the
start of a process took 338, 200, 191, 184, 181, 182, 181, 180, 179, 178 us per
sweep and stayed at 177-178 from there on (about 40 ms of load); a timing taken
inside that ramp favours whatever is measured later.  Every timed region of the
drivers is therefore preceded by at least WARM_MS of the same work, as the
reference's drivers skip their first SKIP_TIMES of RUN_TIMES repetitions
(common/param.hpp:3-9)."""
import time

WARM_MS = 100.0


def warm_chunk(sweeps, fuse=1):
    """Sweeps per warm-up call: an even count of at least `sweeps` that is a
    multiple of 2 * fuse, so that fused runs warm up on full launches (a run
    keeps its result in buffer sweeps % 2, so it splits an odd launch count)."""
    unit = 2 * max(int(fuse or 1), 1)
    return max(unit, (sweeps + unit - 1) // unit * unit)


def warm_up(launch, min_ms=WARM_MS, max_calls=100000):
    """Call launch() (GPU work on the current stream) and wait for it, again and
    again, until min_ms of wall time has passed; returns the number of calls (at
    least one).  Each call should keep the GPU busy for a millisecond or more, so
    that the waits leave no idle gaps worth the name."""
    import torch

    t0 = time.perf_counter()
    n = 0
    while True:
        launch()
        n += 1
        torch.cuda.synchronize()
        if (time.perf_counter() - t0) * 1e3 >= min_ms or n >= max_calls:
            return n
