"""Independent assembly of the linear discontinuous-Galerkin transport system
on tetrahedra, used to pin the oracle (not a restatement of its formulas).

For a direction Omega and a cell K with barycentric basis lambda_i, the
upwind DG weak form of  Omega . grad psi + sigma psi = q  is, for each test
function lambda_i,

  sum_j psi_j [ -int_K lambda_j Omega.grad(lambda_i) dV
                + sigma int_K lambda_i lambda_j dV
                + sum_{faces f, Omega.n_f > 0} (Omega.n_f) int_f lambda_i lambda_j dS ]
  = int_K q lambda_i dV - sum_{faces f, Omega.n_f < 0} (Omega.n_f) int_f psi_up lambda_i dS

Every integral here is evaluated by a quadrature rule exact for quadratics
(4 points in the tetrahedron, 3 in each triangle) with the basis functions
obtained by inverting the vertex matrix, so no closed-form DG entry is
reused.
"""
import numpy as np

_TET_A, _TET_B = 0.5854101966249685, 0.1381966011250105
TET_BARY = np.array([[_TET_A, _TET_B, _TET_B, _TET_B],
                     [_TET_B, _TET_A, _TET_B, _TET_B],
                     [_TET_B, _TET_B, _TET_A, _TET_B],
                     [_TET_B, _TET_B, _TET_B, _TET_A]])
TRI_BARY = np.array([[2 / 3, 1 / 6, 1 / 6], [1 / 6, 2 / 3, 1 / 6],
                     [1 / 6, 1 / 6, 2 / 3]])


def basis(P):
    """Barycentric basis of the tet with vertices P (4x3): returns
    (coef 4x4, grads 4x3) with lambda_i(x) = coef[i] . [x, y, z, 1]."""
    A = np.hstack([P, np.ones((4, 1))])        # rows: [x_j, 1]
    coef = np.linalg.inv(A).T                  # lambda_i(x_j) = delta_ij
    return coef, coef[:, :3]


def face_data(P):
    """Outward unit normals (4x3) and areas of the faces opposite each
    vertex."""
    n = np.zeros((4, 3))
    area = np.zeros(4)
    for k in range(4):
        f = [j for j in range(4) if j != k]
        c = np.cross(P[f[1]] - P[f[0]], P[f[2]] - P[f[0]])
        area[k] = 0.5 * np.linalg.norm(c)
        u = c / np.linalg.norm(c)
        if np.dot(u, P[k] - P[f[0]]) > 0:
            u = -u
        n[k] = u
    return n, area


def local_system(P, omega, sigma, q_nodes, upwind):
    """4x4 matrix and right-hand side of one cell.  upwind[k] (length 4,
    entry k ignored) is the incoming trace at the cell's vertices through
    face k, or None for no inflow through that face."""
    coef, grad = basis(P)
    vol = abs(np.linalg.det(np.hstack([P, np.ones((4, 1))]))) / 6.0
    n, area = face_data(P)
    pts = TET_BARY @ P
    lam = np.hstack([pts, np.ones((4, 1))]) @ coef.T   # [qp][i]
    w = vol / 4.0
    M = np.zeros((4, 4))
    rhs = np.zeros(4)
    od = grad @ omega                                    # Omega.grad lambda_i
    for qp in range(4):
        li = lam[qp]
        M += w * (-np.outer(od, li) + sigma * np.outer(li, li))
        rhs += w * li * (li @ q_nodes)
    for k in range(4):
        on = float(omega @ n[k])
        f = [j for j in range(4) if j != k]
        fpts = TRI_BARY @ P[f]
        flam = np.hstack([fpts, np.ones((3, 1))]) @ coef.T
        fw = area[k] / 3.0
        if on > 0:
            for qp in range(3):
                M += on * fw * np.outer(flam[qp], flam[qp])
        elif on < 0 and upwind[k] is not None:
            for qp in range(3):
                rhs -= on * fw * flam[qp] * (flam[qp] @ upwind[k])
    return M, rhs


def global_system(coords, cell_nodes, adj, omega, sigma, q_nodes):
    """(A, b) of the global DG system A psi = b of one direction, psi
    ordered [c][v] (vacuum inflow on the boundary)."""
    nc = cell_nodes.shape[0]
    A = np.zeros((4 * nc, 4 * nc))
    b = np.zeros(4 * nc)
    for c in range(nc):
        P = coords[cell_nodes[c]]
        coef, _ = basis(P)
        n, area = face_data(P)
        M, rhs = local_system(P, omega, sigma[c], q_nodes[c],
                              [None] * 4)
        A[4 * c:4 * c + 4, 4 * c:4 * c + 4] += M
        b[4 * c:4 * c + 4] += rhs
        for k in range(4):
            u = adj[c, k]
            on = float(omega @ n[k])
            if u < 0 or not on < 0:
                continue
            f = [j for j in range(4) if j != k]
            fpts = TRI_BARY @ P[f]
            flam = np.hstack([fpts, np.ones((3, 1))]) @ coef.T
            fw = area[k] / 3.0
            # upwind trace at cell vertex j = psi[u][vertex of u at node j]
            for j in f:
                vu = int(np.nonzero(cell_nodes[u] == cell_nodes[c, j])[0][0])
                for qp in range(3):
                    A[4 * c:4 * c + 4, 4 * u + vu] += \
                        on * fw * flam[qp] * flam[qp][j]
    return A, b


def global_sweep_solution(coords, cell_nodes, adj, omega, sigma, q_nodes):
    """psi[c][v] for one direction by a dense solve of the global system."""
    A, b = global_system(coords, cell_nodes, adj, omega, sigma, q_nodes)
    return np.linalg.solve(A, b).reshape(cell_nodes.shape[0], 4)
