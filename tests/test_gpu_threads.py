"""Distinct contexts used from distinct host threads at the same time give
the same bits as the same calls made one after another (the library keeps
no per-call state shared between threads: the batch launch recorder and the
kernel timers are per thread, the per-context workspace maps are guarded).
ctypes releases the GIL during the library calls, so the threads overlap."""
import threading

import numpy as np
import pytest

import igr_sensitivities_amd as igr
from igr_sensitivities_amd import configs

pytestmark = pytest.mark.gpu


def _mesh(N):
    return igr.CartesianMesh2D(N, N, 1.0, 1.0, igr.DGBasis(3))


def _gradient(c, q0, dt):
    c.set_state(q0)
    J, qb, gb, ab = c.gradient(0, 4, dt, 1.4, 2e-3, 15, "pcg")
    return J, qb, gb, ab


def _batch(ctxs, q0s, dt):
    runs = []
    for c, q in zip(ctxs, q0s):
        c.set_state(q)
        runs.append(dict(ctx=c, s0=np.stack([q * 0.01]), dt=dt, T=3 * dt, gamma=1.4, alpha=1e-3, sgamma=0.0))
    out = igr.run_sensitivity_batch(runs, n_iter=10, solver="pcg")
    return [(c.get_state()[0], o[0]) for c, o in zip(ctxs, out)]


def _steps(c, q0, dt):
    c.set_state(q0)
    c.run_steps(5, dt, 1.4, 1e-3, 20, "pcg")
    return c.get_state()


def test_two_host_threads_bit_equal_to_serial():
    print("[case] two host threads: gradient + batch + plain runs", flush=True)
    N = 10
    q0 = configs.smooth_random2(3, N, N, seed=21, amp=0.2)
    qs = [configs.smooth_random2(3, N, N, seed=30 + k, amp=0.2) for k in range(3)]
    dt = 0.3 * (1 / N) / (7 * configs.max_wavespeed2(q0, 1.4))
    cg, cs = igr.Context(_mesh(N)), igr.Context(_mesh(N))
    cb = [igr.Context(_mesh(N)) for _ in range(3)]
    ref_g = _gradient(cg, q0, dt)
    ref_b = _batch(cb, qs, dt)
    ref_s = _steps(cs, q0, dt)

    res = {}
    errs = []

    def run(name, fn):
        try:
            for _ in range(3):
                res[name] = fn()
        except Exception as e:   # reported by the main thread
            errs.append((name, e))

    ths = [threading.Thread(target=run, args=("g", lambda: _gradient(cg, q0, dt))),
           threading.Thread(target=run, args=("b", lambda: _batch(cb, qs, dt))),
           threading.Thread(target=run, args=("s", lambda: _steps(cs, q0, dt)))]
    for t in ths:
        t.start()
    for t in ths:
        t.join(timeout=100)
    assert not any(t.is_alive() for t in ths)
    assert not errs, errs
    J, qb, gb, ab = res["g"]
    assert J == ref_g[0] and gb == ref_g[2] and ab == ref_g[3]
    np.testing.assert_array_equal(qb, ref_g[1])
    for (q, s), (qr, sr) in zip(res["b"], ref_b):
        np.testing.assert_array_equal(q, qr)
        np.testing.assert_array_equal(s, sr)
    np.testing.assert_array_equal(res["s"][0], ref_s[0])
    np.testing.assert_array_equal(res["s"][1], ref_s[1])
    for c in [cg, cs] + cb:
        c.close()
