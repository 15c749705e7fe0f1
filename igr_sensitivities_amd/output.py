"""Output formats of the reference's paper scripts, for fields copied back
from the device (host-side numpy; nothing here computes the path):

* DG fields flattened for line plots and CSV: each element's polynomial
  evaluated on equispaced points, with a NaN row between elements
  (flatten_dg, IGR_adjoints_1d/paper_scripts/generate_shift_sensitivity_data.jl:104-127,
  with its Lagrange interpolation matrix, :61-93);
* conservative -> primitive variables (:134-138);
* tensor-product 2D fields flattened to the logical global grid used for
  heatmaps, with its node coordinates made strictly increasing
  (IGR_adjoints_2d/paper_scripts/generate_sedov_heatmap_data.jl:126-136).

Array layouts are this package's: a 1D field is (N_e, n_p), a 2D field
(N_ey, N_ex, n_p, n_p) indexed [ey, ex, j, i].
"""
import numpy as np

from .configs import coords1, coords2, lgl


def lagrange_interpolation_matrix(xi_from, xi_to):
    """V with V @ f(xi_from) = the degree n-1 interpolant of f at xi_to, in the
    barycentric (second) form of Berrut and Trefethen (SIAM Review 46, 2004);
    a target within 1e-14 of a node takes that node's value."""
    xi_from = np.asarray(xi_from, dtype=np.float64)
    xi_to = np.asarray(xi_to, dtype=np.float64)
    diff = xi_from[:, None] - xi_from[None, :]
    np.fill_diagonal(diff, 1.0)
    lam = 1.0 / np.prod(diff, axis=0)          # lambda_j = 1 / prod_{i != j} (x_j - x_i)
    V = np.zeros((xi_to.size, xi_from.size))
    for k, x in enumerate(xi_to):
        d = x - xi_from
        hit = np.nonzero(np.abs(d) < 1e-14)[0]
        if hit.size:
            V[k, hit[0]] = 1.0
        else:
            t = lam / d
            V[k] = t / np.sum(t)
    return V


def flatten_dg(q, p, L=1.0, n_sub=10):
    """(x_out, q_out): a 1D DG field (N_e, n_p) on n_sub equispaced points per
    element, NaN between elements so a plot shows each element's polynomial
    without joining jumps."""
    q = np.asarray(q, dtype=np.float64)
    Ne, npn = q.shape
    x = coords1(p, Ne, L)
    xi, _ = lgl(p)
    xi_fine = np.linspace(-1.0, 1.0, n_sub)
    V = lagrange_interpolation_matrix(xi, xi_fine)
    xs, qs = [], []
    for e in range(Ne):
        mid, half = (x[e, 0] + x[e, -1]) / 2, (x[e, -1] - x[e, 0]) / 2
        xs.append(mid + half * xi_fine)
        qs.append(V @ q[e])
        if e < Ne - 1:
            xs.append([np.nan])
            qs.append([np.nan])
    return np.concatenate(xs), np.concatenate(qs)


def conservative_to_primitive(mu, rho, E, gamma):
    """(u, rho, p) of a 1D state."""
    u = mu / rho
    return u, rho, (gamma - 1) * (E - mu ** 2 / (2 * rho))


def conservative_to_primitive2(mx, my, rho, E, gamma):
    """(u_x, u_y, rho, p) of a 2D state (mechanical pressure)."""
    return mx / rho, my / rho, rho, (gamma - 1) * (E - (mx ** 2 + my ** 2) / (2 * rho))


def flatten2d(field):
    """The logical global grid Q[ex*n_p + i, ey*n_p + j] = field[ey, ex, j, i]."""
    f = np.asarray(field)
    Ney, Nex, npn, _ = f.shape
    return f.transpose(1, 3, 0, 2).reshape(Nex * npn, Ney * npn)


def logical_grid(p, Nex, Ney, Lx=1.0, Ly=1.0):
    """(xg, yg): node coordinates along the logical grid's axes; the face
    nodes shared by neighbouring elements are separated by a ramp of total
    height 1e-9 so both axes increase strictly (heatmap renderers need it)."""
    x, y = coords2(p, Nex, Ney, Lx, Ly)
    xg = x[0, :, 0, :].reshape(-1).copy()
    yg = y[:, 0, :, 0].reshape(-1).copy()
    xg += np.linspace(0.0, 1e-9, xg.size)
    yg += np.linspace(0.0, 1e-9, yg.size)
    return xg, yg


def write_csv(path, columns):
    """Columns (name -> 1D array of one length) as CSV with a header row."""
    names = list(columns)
    data = np.column_stack([np.asarray(columns[n], dtype=np.float64) for n in names])
    np.savetxt(path, data, delimiter=",", header=",".join(names), comments="", fmt="%.16e")
