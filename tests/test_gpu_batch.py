"""Batched ensembles (igr_batch_run_sensitivity): every instance of a mixed
1D/2D batch of uneven sizes, step counts, gamma, alpha and tangent counts
must equal, bit for bit, its own igr_run_sensitivity call (the reference
runs run_sensitivity once per instance, IGR_adjoints_2d/sensitivity.jl:353,
IGR_adjoints_1d/sensitivity.jl:475); the single runs are in turn pinned to
the oracle by test_gpu_tangent.py."""
import numpy as np
import pytest

import igr_sensitivities_amd as igr
from igr_sensitivities_amd import configs

pytestmark = pytest.mark.gpu


def make_runs(solver_alpha, seed=11):
    rng = np.random.default_rng(seed)
    runs = []
    # 2D: uneven meshes (12 is not a multiple of the 4x4 tile), step counts and tangent counts
    for k, (n, nt) in enumerate([(8, 2), (12, 1), (16, 0), (8, 3)]):
        mesh = igr.CartesianMesh2D(n, n, 1.0, 1.0, igr.DGBasis(3))
        q0 = configs.smooth_random2(3, n, n, seed=100 + k)
        gam = float(rng.uniform(1.2, 1.67))
        dt = 0.3 * (1.0 / n) / (7 * configs.max_wavespeed2(q0, gam))
        s0 = rng.standard_normal((nt, 4) + mesh.shape) * 0.1 if nt else None
        # instances 1 and 2 run the PCG x/r update on fewer workgroups than
        # tiles (9 tiles on 4, 16 on 5: uneven tiles per workgroup)
        runs.append(dict(mesh=mesh, q0=q0, s0=s0, dt=dt, T=(3 + k) * dt * 0.97, gamma=gam,
                         alpha=solver_alpha[1] * (1.0 if k != 2 else 0.5), groups=(0, 4, 5, 0)[k]))
    # 1D: N_e = 64 (256 nodes, smaller workgroup) and 256/320 (1024-thread workgroups)
    for k, (ne, nt) in enumerate([(256, 2), (64, 1), (320, 4)]):
        mesh = igr.PeriodicMesh1D(ne, 1.0, igr.DGBasis(3))
        q0 = configs.smooth_random1(3, ne, seed=200 + k)
        gam = float(rng.uniform(1.2, 1.67))
        dt = 0.1 * (1.0 / ne) / (7 * configs.max_wavespeed1(q0, gam))
        s0 = rng.standard_normal((nt, 3) + mesh.shape) * 0.1
        runs.append(dict(mesh=mesh, q0=q0, s0=s0, dt=dt, T=(4 + k) * dt * 0.93, gamma=gam,
                         alpha=solver_alpha[1], sgamma=0.5 if k == 1 else 0.0))
    return runs


@pytest.mark.parametrize("solver_alpha", [("pcg", 2e-3), ("jacobi", 2e-3), ("chebyshev", 2e-3), ("pcg", 0.0)])
def test_batch_equals_single_runs(solver_alpha):
    solver = solver_alpha[0]
    print(f"[case] batch vs single, {solver_alpha}", flush=True)
    runs = make_runs(solver_alpha)
    n_iter = 12
    # single runs
    ref = []
    for r in runs:
        c = igr.Context(r["mesh"])
        if r.get("groups"):
            c.set_option(igr.core.OPT_GROUPS, r["groups"])
        c.set_state(r["q0"])
        s0 = r["s0"] if r["s0"] is not None else np.empty((0, c.nv) + c.shape)
        if s0.shape[0]:
            s, vs, ns = c.run_sensitivity(s0, r["dt"], r["T"], r["gamma"], r["alpha"], n_iter, solver,
                                          2, r.get("sgamma", 0.0))
        else:
            ns = c.run_forward(r["dt"], r["T"], r["gamma"], r["alpha"], n_iter, solver) if r["mesh"].dim == 2 \
                else None
            s = vs = None
        q, S = c.get_state()
        ref.append((q, S, s, vs, ns))
        c.close()
    # one batch
    ctxs = []
    for r in runs:
        c = igr.Context(r["mesh"])
        if r.get("groups"):
            c.set_option(igr.core.OPT_GROUPS, r["groups"])
        c.set_state(r["q0"])
        ctxs.append(c)
    out = igr.run_sensitivity_batch([dict(r, ctx=c) for r, c in zip(runs, ctxs)], n_iter=n_iter, solver=solver)
    for k, (r, c) in enumerate(zip(runs, ctxs)):
        q, S = c.get_state()
        qr, Sr, sr, vsr, nsr = ref[k]
        np.testing.assert_array_equal(q, qr, err_msg=f"instance {k} state")
        np.testing.assert_array_equal(S, Sr, err_msg=f"instance {k} Sigma")
        s, vs, ns = out[k]
        if sr is not None:
            np.testing.assert_array_equal(s, sr, err_msg=f"instance {k} tangents")
            np.testing.assert_array_equal(vs, vsr, err_msg=f"instance {k} varsigma")
        if nsr is not None:
            assert ns == nsr
        c.close()


def test_batch_uneven_step_counts_drop_out():
    """Instances with fewer steps stop while the others go on."""
    print("[case] batch uneven steps", flush=True)
    mesh = igr.CartesianMesh2D(8, 8, 1.0, 1.0, igr.DGBasis(3))
    q0 = configs.taylor_green2(3, 8, 8, A=0.2)
    dt = 0.3 * (1 / 8) / (7 * configs.max_wavespeed2(q0, 1.4))
    ctxs, runs = [], []
    for k in range(3):
        c = igr.Context(mesh)
        c.set_state(q0)
        ctxs.append(c)
        runs.append(dict(ctx=c, s0=None, dt=dt, T=(1.5 + 3 * k) * dt, gamma=1.4, alpha=1e-3))
    out = igr.run_sensitivity_batch(runs, n_iter=10, solver="pcg")
    assert [o[2] for o in out] == [2, 5, 8]
    for k, c in enumerate(ctxs):
        d = igr.Context(mesh)
        d.set_state(q0)
        d.run_forward(dt, (1.5 + 3 * k) * dt, 1.4, 1e-3, 10, "pcg")
        np.testing.assert_array_equal(c.get_state()[0], d.get_state()[0])


def test_batch_errors():
    print("[case] batch errors", flush=True)
    mesh = igr.CartesianMesh2D(8, 8, 1.0, 1.0, igr.DGBasis(3))
    c = igr.Context(mesh)
    c.set_state(configs.taylor_green2(3, 8, 8))
    r = dict(ctx=c, s0=None, dt=1e-3, T=2e-3, gamma=1.4, alpha=1e-3)
    with pytest.raises(ValueError):
        igr.run_sensitivity_batch([r, r])          # the same context twice
    with pytest.raises(ValueError):
        igr.run_sensitivity_batch([r], solver="gmres")
