"""The two larger workloads at their stated sizes, on the GPU:

- the blast beside the Oseen vortex pair at 256 x 256, p = 3, PCG-60: two
  SSP-RK3 steps against the oracle (1e-12), the gradient of the disc kinetic
  energy over two steps against the oracle's dual-number JVP (the dot-product
  identity, 1e-9), and the full 400-step gradient in gamma and alpha against
  central differences of the GPU forward map;
- the 64-instance ensemble (1D N_e 256..4096 and 2D 32^2..128^2, 4 tangents
  each, damped Jacobi-50) as one batched run, bit-equal to the instances run
  one by one, with its smallest 1D and 2D instances against the oracle's
  run_sensitivity (IGR_adjoints_1d/sensitivity.jl:475, IGR_adjoints_2d/
  sensitivity.jl:353).

The oracle runs its OpenMP build here: the same loops, bit-equal to the
serial build (tests/test_oracle_omp.py)."""
import numpy as np
import pytest

import igr_sensitivities_amd as igr
from igr_sensitivities_amd import configs

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def close(a, b, tol):
    assert abs(a - b) <= tol * max(abs(a), abs(b), 1e-300), (a, b, abs(a - b) / max(abs(a), abs(b), 1e-300))


@pytest.fixture(scope="module")
def c5():
    return configs.config5()


def _ctx256():
    print("[case] blast-vortex 256x256", flush=True)
    return igr.Context(igr.CartesianMesh2D(256, 256, 1.0, 1.0, igr.DGBasis(3)))


def test_c5_two_steps_vs_oracle(orc, c5):
    c = _ctx256()
    assert c.get_option(igr.core.OPT_TILE) == 804      # the 256^2 default shapes
    c.set_state(c5["q"])
    c.run_steps(2, c5["dt"], c5["gamma"], c5["alpha"], 60, "pcg")
    q, S = c.get_state()
    c.close()
    g = orc.Grid2D(3, 256, 256)
    qr, Sr = g.run_steps(c5["gamma"], c5["alpha"], 60, orc.PCG, 2, c5["dt"], c5["q"], omp=True)
    assert rel(q, qr) < 1e-12
    assert rel(S, Sr) < 1e-10     # a 60-iteration PCG iterate: the dot products' order shows at ~1e-11


def test_c5_two_step_gradient_vs_oracle_jvp(orc, c5):
    c = _ctx256()
    q0, gam, a, dt = c5["q"], c5["gamma"], c5["alpha"], c5["dt"]
    c.set_state(q0)
    J, qb, gb, ab = c.gradient(6, 2, dt, gam, a, 60, "pcg", prm=c5["disc"])
    c.close()
    u = np.random.default_rng(82).standard_normal(q0.shape) * (q0.std(axis=(1, 2, 3, 4), keepdims=True) + 1e-2)
    g = orc.Grid2D(3, 256, 256)
    Jr, dJ, _, _ = g.jvp_objective_run(gam, 0.3, a, 0.2 * a, 60, orc.PCG, 2, dt, 6, q0, u, c5["disc"], omp=True)
    close(J, Jr, 1e-12)
    close(dJ, np.vdot(qb, u) + 0.3 * gb + 0.2 * a * ab, 1e-9)


def test_c5_full_gradient_vs_central_differences(c5):
    """All 400 steps under a budget of 40 checkpoints (binomial recomputation);
    dJ/dgamma and dJ/dalpha against central differences of the GPU map with a
    relative step of 1e-6 (truncation ~1e-12, rounding of J ~1e-10)."""
    c = _ctx256()
    q0, gam, a, dt, ns = c5["q"], c5["gamma"], c5["alpha"], c5["dt"], c5["n_steps"]
    obj = igr.OBJECTIVES["disc_kinetic_energy"]
    c.set_state(q0)
    J, _, gb, ab = c.gradient(obj, ns, dt, gam, a, 60, "pcg", prm=c5["disc"], ckpt_slots=40)

    def Jat(gg, aa):
        c.set_state(q0)
        c.run_steps(ns, dt, gg, aa, 60, "pcg")
        return c.objective(obj, gg, c5["disc"])

    h = 1e-6
    close(Jat(gam, a), J, 1e-13)
    fd_g = (Jat(gam * (1 + h), a) - Jat(gam * (1 - h), a)) / (2 * h * gam)
    fd_a = (Jat(gam, a * (1 + h)) - Jat(gam, a * (1 - h))) / (2 * h * a)
    c.close()
    igr.release_cache()
    eg = abs(fd_g - gb) / abs(gb)
    ea = abs(fd_a - ab) / abs(ab)
    print(f"[c5 400 steps] dJ/dgamma {gb:.12e} vs FD {fd_g:.12e} (rel {eg:.2e}); "
          f"dJ/dalpha {ab:.12e} vs FD {fd_a:.12e} (rel {ea:.2e})", flush=True)
    assert eg < 1e-6 and ea < 1e-6


def test_c4_full_batch_bit_equal_and_smallest_vs_oracle(orc):
    c4 = configs.config4()
    runs = c4["runs"]
    print(f"[case] ensemble of {len(runs)} instances", flush=True)
    ctxs = []
    for r in runs:
        mesh = (igr.PeriodicMesh1D(r["N"], 1.0, igr.DGBasis(3)) if r["dim"] == 1
                else igr.CartesianMesh2D(r["N"], r["N"], 1.0, 1.0, igr.DGBasis(3)))
        ctxs.append(igr.Context(mesh))
    for r, c in zip(runs, ctxs):
        c.set_state(r["q0"])
    out = igr.run_sensitivity_batch([dict(r, ctx=c) for r, c in zip(runs, ctxs)], n_iter=c4["n_iter"],
                                    solver=c4["solver"])
    qT = [c.get_state() for c in ctxs]
    for k, (r, c) in enumerate(zip(runs, ctxs)):
        c.set_state(r["q0"])
        s, vs, ns = c.run_sensitivity(r["s0"], r["dt"], r["T"], r["gamma"], r["alpha"], c4["n_iter"],
                                      c4["solver"])
        assert ns == out[k][2] == r["n_steps"]
        np.testing.assert_array_equal(out[k][0], s)
        np.testing.assert_array_equal(out[k][1], vs)
        q, S = c.get_state()
        np.testing.assert_array_equal(qT[k][0], q)
    for c in ctxs:
        c.close()
    # the smallest 1D and 2D instances, every tangent, against the oracle
    for dim in (1, 2):
        k = min((k for k, r in enumerate(runs) if r["dim"] == dim), key=lambda k: (runs[k]["N"], runs[k]["n_steps"]))
        r = runs[k]
        g = orc.Grid1D(3, r["N"]) if dim == 1 else orc.Grid2D(3, r["N"], r["N"])
        for t in range(r["s0"].shape[0]):
            if dim == 1:
                qr, Sr, sr, vr, nsr = g.run_sensitivity(r["gamma"], r["alpha"], c4["n_iter"], orc.JACOBI, r["dt"],
                                                        r["T"], r["q0"], r["s0"][t], form=2)
            else:
                qr, Sr, sr, vr, nsr = g.run_sensitivity(r["gamma"], r["alpha"], c4["n_iter"], orc.JACOBI, r["dt"],
                                                        r["T"], r["q0"], r["s0"][t])
            assert nsr == out[k][2]
            assert rel(qT[k][0], qr) < 1e-10, (dim, t)
            assert rel(out[k][0][t], sr) < 1e-9, (dim, t)
            assert rel(out[k][1][t], vr) < 1e-8, (dim, t)
