"""Output formats of the paper scripts (igr_sensitivities_amd.output): the
Lagrange interpolation is exact on polynomials of the basis degree and the
identity at the nodes; flatten_dg keeps the reference's NaN-separated layout
(IGR_adjoints_1d/paper_scripts/generate_shift_sensitivity_data.jl:104-127);
flatten2d follows Q[(ex-1)n_p + i, (ey-1)n_p + j] = field[i,j,ex,ey]
(IGR_adjoints_2d/paper_scripts/generate_sedov_heatmap_data.jl:126-132)."""
import numpy as np
import pytest

from igr_sensitivities_amd import configs, output


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_interpolation_exact_on_polynomials(p):
    xi, _ = configs.lgl(p)
    V = output.lagrange_interpolation_matrix(xi, xi)
    np.testing.assert_array_equal(V, np.eye(p + 1))
    t = np.linspace(-1, 1, 17)
    V = output.lagrange_interpolation_matrix(xi, t)
    for k in range(p + 1):
        assert np.max(np.abs(V @ xi ** k - t ** k)) < 1e-13
    # one degree more is not reproduced
    assert np.max(np.abs(V @ xi ** (p + 1) - t ** (p + 1))) > 1e-3


def test_flatten_dg_layout_and_values():
    p, Ne, n_sub = 3, 5, 7
    x = configs.coords1(p, Ne)
    f = lambda z: 1 + 2 * z - 3 * z ** 2 + z ** 3     # degree 3: exact inside each element
    xo, qo = output.flatten_dg(f(x), p, n_sub=n_sub)
    assert xo.shape == qo.shape == (Ne * n_sub + Ne - 1,)
    nan = np.isnan(xo)
    assert nan.sum() == Ne - 1 and np.array_equal(nan, np.isnan(qo))
    assert np.all(nan[n_sub::n_sub + 1])
    assert np.max(np.abs(qo[~nan] - f(xo[~nan]))) < 1e-12
    # element ends land on the element's end nodes
    assert xo[0] == pytest.approx(x[0, 0]) and xo[n_sub - 1] == pytest.approx(x[0, -1])


def test_primitive_variables():
    q = configs.smooth_random1(3, 4, seed=2, amp=0.3)
    u, r, pr = output.conservative_to_primitive(q[0], q[1], q[2], 1.4)
    np.testing.assert_allclose(pr / 0.4 + 0.5 * r * u ** 2, q[2], rtol=1e-14)
    q2 = configs.smooth_random2(3, 3, 2, seed=2, amp=0.3)
    ux, uy, r2, p2 = output.conservative_to_primitive2(*q2, 1.4)
    np.testing.assert_allclose(p2 / 0.4 + 0.5 * r2 * (ux ** 2 + uy ** 2), q2[3], rtol=1e-14)


def test_flatten2d_and_grid():
    p, Nex, Ney = 2, 4, 3
    npn = p + 1
    rng = np.random.default_rng(0)
    field = rng.standard_normal((Ney, Nex, npn, npn))
    Q = output.flatten2d(field)
    assert Q.shape == (Nex * npn, Ney * npn)
    for ey in range(Ney):
        for ex in range(Nex):
            for j in range(npn):
                for i in range(npn):
                    assert Q[ex * npn + i, ey * npn + j] == field[ey, ex, j, i]
    xg, yg = output.logical_grid(p, Nex, Ney, 1.0, 0.5)
    assert xg.shape == (Nex * npn,) and yg.shape == (Ney * npn,)
    assert np.all(np.diff(xg) > 0) and np.all(np.diff(yg) > 0)
    x, y = configs.coords2(p, Nex, Ney, 1.0, 0.5)
    assert np.max(np.abs(xg - x[0, :, 0, :].ravel())) <= 1.01e-9


def test_write_csv(tmp_path):
    path = tmp_path / "out.csv"
    xo, qo = output.flatten_dg(configs.coords1(2, 3) ** 2, 2, n_sub=4)
    output.write_csv(path, {"x": xo, "q": qo})
    lines = path.read_text().splitlines()
    assert lines[0] == "x,q" and len(lines) == xo.size + 1
    back = np.genfromtxt(path, delimiter=",", skip_header=1)
    np.testing.assert_array_equal(np.isnan(back[:, 0]), np.isnan(xo))
    np.testing.assert_array_equal(back[~np.isnan(xo)], np.column_stack([xo, qo])[~np.isnan(xo)])
