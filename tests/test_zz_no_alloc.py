"""igr_solve_elliptic_noalloc (no device allocation; work arrays in the
library's fixed scratch array) against igr_solve_elliptic on the same inputs:
bit-equal results, on two calls in a row with the second on a smaller mesh,
for every solver; its limits and its domain flag."""
import ctypes as C

import numpy as np
import pytest

import igr_sensitivities_amd as igr
from igr_sensitivities_amd import _lib, configs

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Dev:
    """Device buffers through the HIP runtime (the test allocates; the entry
    point under test does not)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.dtype in (np.float64, np.int32)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 8))) == 0
        self.bufs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), H2D) == 0
        return p

    def get(self, p, shape, dtype=np.float64):
        out = np.empty(shape, dtype=dtype)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def _inputs(nx, ny, seed):
    q = configs.smooth_random2(3, nx, ny, seed=seed, amp=0.3)
    s0 = 0.01 * np.random.default_rng(seed + 1).standard_normal(q.shape[1:])
    return q, s0


def _noalloc(dev, nx, ny, alpha, n_iter, solver, q, s0, flag=True):
    L = _lib.lib()
    d_mx, d_my, d_rho, d_s = dev.put(q[0]), dev.put(q[1]), dev.put(q[2]), dev.put(s0)
    d_flag = dev.put(np.full(2, 7, dtype=np.int32)) if flag else None
    rc = L.igr_solve_elliptic_noalloc(3, nx, ny, 1.0, 1.0, alpha, n_iter, solver, d_mx, d_my, d_rho, d_s,
                                      d_flag, None)
    S = dev.get(d_s, s0.shape) if rc == 0 else None
    f = int(dev.get(d_flag, (2,), np.int32)[0]) if (flag and rc == 0) else None
    return rc, S, f


def _context(nx, ny, alpha, n_iter, solver, q, s0):
    c = igr.Context(igr.CartesianMesh2D(nx, ny, 1.0, 1.0, igr.DGBasis(3)))
    c.set_state(q, s0)
    c.solve_elliptic(alpha, n_iter, solver)
    S = c.get_state()[1]
    c.close()
    return S


@pytest.mark.parametrize("solver", [0, 1, 2])
def test_noalloc_matches_solve_elliptic_twice(solver):
    dev = Dev()
    try:
        for nx, ny, seed in ((24, 20, 50), (10, 6, 51)):      # the second call on a smaller mesh
            print(f"[case] noalloc solver={solver} {nx}x{ny}", flush=True)
            q, s0 = _inputs(nx, ny, seed)
            a = configs.alpha_of(1 / nx, 3)
            rc, S, f = _noalloc(dev, nx, ny, a, 25, solver, q, s0)
            assert rc == 0 and f == 0
            np.testing.assert_array_equal(S, _context(nx, ny, a, 25, solver, q, s0))
    finally:
        dev.free()


def test_noalloc_at_its_size_limit_and_beyond():
    dev = Dev()
    try:
        print("[case] noalloc 128x128 (the limit)", flush=True)
        cfg = configs.config2(N=128)
        q, s0 = cfg["q"], np.zeros(cfg["q"].shape[1:])
        rc, S, f = _noalloc(dev, 128, 128, cfg["alpha"], 60, 0, q, s0)
        assert rc == 0 and f == 0
        np.testing.assert_array_equal(S, _context(128, 128, cfg["alpha"], 60, 0, q, s0))
        L = _lib.lib()
        # one element row more than the scratch array holds, an unknown solver,
        # too many iterations: refused before anything is enqueued
        big = dev.put(np.zeros(16 * 129 * 128, dtype=np.float64))
        assert L.igr_solve_elliptic_noalloc(3, 129, 128, 1.0, 1.0, 1e-3, 10, 0, big, big, big, big, None,
                                            None) == _lib.IGR_EINVAL
        small = dev.put(np.ones(16 * 4, dtype=np.float64))
        assert L.igr_solve_elliptic_noalloc(3, 2, 2, 1.0, 1.0, 1e-3, 10, 3, small, small, small, small, None,
                                            None) == _lib.IGR_EINVAL
        assert L.igr_solve_elliptic_noalloc(3, 2, 2, 1.0, 1.0, 1e-3, 5000, 0, small, small, small, small, None,
                                            None) == _lib.IGR_EINVAL
    finally:
        dev.free()


def test_noalloc_domain_flag():
    dev = Dev()
    try:
        print("[case] noalloc domain flag", flush=True)
        q, s0 = _inputs(6, 6, 60)
        q[2, 1, 2, 0, 0] = -0.5          # one non-positive density
        rc, _, f = _noalloc(dev, 6, 6, 1e-3, 5, 0, q, s0)
        assert rc == 0 and f == 1
        q, s0 = _inputs(6, 6, 61)        # the next call clears it
        rc, _, f = _noalloc(dev, 6, 6, 1e-3, 5, 0, q, s0)
        assert rc == 0 and f == 0
    finally:
        dev.free()
