"""The OpenMP build of the oracle (which times the multi-core CPU baseline)
gives the same bits as the serial build that checks the GPU path: its
parallel loops write disjoint nodes and keep each node's order of additions."""
import numpy as np
import pytest

from igr_sensitivities_amd import configs
from oracle import oracle as O


@pytest.mark.parametrize("Nex,Ney,solver", [(6, 5, O.PCG), (5, 6, O.JACOBI), (4, 4, O.CHEBYSHEV),
                                            (1, 3, O.PCG), (3, 1, O.PCG)])
def test_omp_build_bit_equal(Nex, Ney, solver):
    g = O.Grid2D(3, Nex, Ney)
    q0 = configs.smooth_random2(3, Nex, Ney, seed=11, amp=0.2)
    dt = 0.3 * (1 / max(Nex, Ney)) / (7 * configs.max_wavespeed2(q0, 1.4))
    q1, S1 = g.run_steps(1.4, 2e-3, 9, solver, 3, dt, q0)
    q2, S2, threads = O.run_steps2_omp(g, 1.4, 2e-3, 9, solver, 3, dt, q0)
    assert threads >= 1
    np.testing.assert_array_equal(q2, q1)
    np.testing.assert_array_equal(S2, S1)
