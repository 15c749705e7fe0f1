"""Tile-shape timing report (2D, p = 3): for each mesh size, every SIP-family
shape and every hyperbolic-family shape is timed on a smooth seeded state,
and the autotuner's pick is printed.  One JSON line per mesh size.

    python3 -m igr_sensitivities_amd.tune [N ...]
    python3 -m igr_sensitivities_amd.tune --reverse [N ...]

With --reverse, the gradient (every step kept, PCG-60) is timed for every
shape of the reverse SIP family instead.  The size table of the library's
default shapes was chosen from these reports.
"""
import json
import sys
import time

from . import core as igr
from . import configs
from .core import OPT_TILE, OPT_HTILE, OPT_RTILE


def report(N, steps=10, reps=100):
    q = configs.smooth_random2(3, N, N, seed=7, amp=0.2)
    gam, al = 1.4, configs.alpha_of(1.0 / N, 3)
    dt = 0.4 * (1.0 / N) / (7 * configs.max_wavespeed2(q, gam))
    mesh = igr.CartesianMesh2D(N, N, 1.0, 1.0, igr.DGBasis(3))
    row = {"N": N}
    for tile in (404, 804, 808):
        c = igr.Context(mesh)
        c.set_option(OPT_TILE, tile)
        c.set_option(OPT_HTILE, 804)
        c.set_state(q)
        c.run_steps(2, dt, gam, al, 60, "pcg")
        c.set_state(q)
        t0 = time.perf_counter()
        c.run_steps(steps, dt, gam, al, 60, "pcg")
        row[f"sip{tile}_step_ms"] = (time.perf_counter() - t0) / steps * 1e3
        row[f"sip{tile}_pcgA_us"] = c.time_kernel(0, gam, al, reps) * 1e3
        row[f"sip{tile}_pcgB_us"] = c.time_kernel(1, gam, al, reps) * 1e3
        c.close()
    for htile in (404, 804):
        c = igr.Context(mesh)
        c.set_option(OPT_TILE, 804)
        c.set_option(OPT_HTILE, htile)
        c.set_state(q)
        row[f"hyp{htile}_us"] = c.time_kernel(2, gam, al, reps) * 1e3
        c.close()
    c = igr.Context(mesh)
    c.set_state(q)
    row["autotune_pick"] = c.autotune()
    row["default"] = (igr.Context(mesh).get_option(OPT_TILE), igr.Context(mesh).get_option(OPT_HTILE))
    c.close()
    return row


def reverse_report(N, steps=None):
    """Gradient time per step for each reverse SIP-family shape (best of two
    calls after one that allocates the workspace); the forward step time and
    the size table's shape are reported beside it."""
    steps = steps or max(10, min(40, 40 * 128 * 128 // (N * N)))
    q = configs.smooth_random2(3, N, N, seed=7, amp=0.2)
    gam, al = 1.4, configs.alpha_of(1.0 / N, 3)
    dt = 0.4 * (1.0 / N) / (7 * configs.max_wavespeed2(q, gam))
    obj = igr.OBJECTIVES["disc_kinetic_energy"]
    c = igr.Context(igr.CartesianMesh2D(N, N, 1.0, 1.0, igr.DGBasis(3)))
    row = {"N": N, "steps": steps, "default": c.get_option(OPT_RTILE)}
    c.set_state(q)
    c.run_steps(2, dt, gam, al, 60, "pcg")
    c.set_state(q)
    t0 = time.perf_counter()
    c.run_steps(steps, dt, gam, al, 60, "pcg")
    row["forward_step_ms"] = (time.perf_counter() - t0) / steps * 1e3
    for rtile in (404, 804, 808):
        c.set_option(OPT_RTILE, rtile)
        best = float("inf")
        for rep in range(3):
            c.set_state(q)
            t0 = time.perf_counter()
            c.gradient(obj, steps, dt, gam, al, 60, "pcg")
            if rep:
                best = min(best, time.perf_counter() - t0)
        row[f"rev{rtile}_gradient_step_ms"] = best / steps * 1e3
    c.close()
    return row


def main(argv):
    rev = "--reverse" in argv
    sizes = [int(a) for a in argv if a != "--reverse"] or [16, 32, 64, 128, 256]
    for N in sizes:
        print(json.dumps(reverse_report(N) if rev else report(N)), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
