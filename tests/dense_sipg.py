"""Brute-force dense assembly of the SIPG Laplace matrix in numpy (tiny meshes only).

An independent statement of the bilinear form (no sum factorization, full 3D/2D quadrature
with QGauss(k+1), numpy's own Legendre machinery for nodes and bases). Used to pin the
oracle's operator and RHS. The form (laplace_operator.cpp:129-265, weak_boundary_conditions.h):
  a(u,v) = sum_K int grad u . grad v
         + sum_F int ( -{du/dn}[v] - [u]{dv/dn} + tau_F [u][v] )      (interior / periodic)
         + sum_{F in Gamma_D} int ( -du/dn v - u dv/dn + 2 tau_F u v )
"""
import itertools

import numpy as np
from numpy.polynomial import legendre as L


def gl_nodes(n):
    if n == 1:
        return np.array([0.5])
    inner = L.Legendre.basis(n - 1).deriv().roots()
    return np.concatenate([[0.0], np.sort(0.5 * (inner + 1.0)), [1.0]])


def gauss(n):
    x, w = L.leggauss(n)
    return 0.5 * (x + 1.0), 0.5 * w


class Basis1D:
    """Lagrange basis on given nodes of [0,1] via a Legendre Vandermonde inverse."""

    def __init__(self, nodes):
        self.n = len(nodes)
        V = L.legvander(2 * np.asarray(nodes) - 1, self.n - 1)
        self.C = np.linalg.inv(V)  # column i = Legendre coefficients of phi_i

    def val(self, x):
        x = np.atleast_1d(x)
        return L.legvander(2 * x - 1, self.n - 1) @ self.C  # [len(x), n]

    def der(self, x):
        x = np.atleast_1d(x)
        out = np.empty((len(x), self.n))
        for i in range(self.n):
            out[:, i] = 2.0 * L.legval(2 * x - 1, L.legder(self.C[:, i]))
        return out


class Mesh:
    def __init__(self, n_cells, lower, upper, deformation=0.0, frequency=1, mapping_degree=1,
                 periodic=(0, 0, 0), bc=(0,) * 6):
        self.nc = np.array(n_cells)
        self.lo = np.array(lower, dtype=float) * np.ones(3)
        self.hi = np.array(upper, dtype=float) * np.ones(3)
        self.h = (self.hi - self.lo) / self.nc
        self.deform, self.freq = deformation, frequency
        self.periodic, self.bc = periodic, bc
        self.mbasis = Basis1D(gl_nodes(mapping_degree + 1)) if deformation != 0 else None
        self.mnodes = gl_nodes(mapping_degree + 1) if deformation != 0 else None
        self.cells = [c for c in itertools.product(*[range(int(n)) for n in self.nc[::-1]])]
        self.cells = [(c[2], c[1], c[0]) for c in self.cells]  # x fastest

    def index(self, cc):
        return cc[0] + self.nc[0] * (cc[1] + self.nc[1] * cc[2])

    def neighbor(self, cc, d, s):
        n = list(cc)
        n[d] += 1 if s else -1
        if n[d] < 0 or n[d] >= self.nc[d]:
            if not self.periodic[d]:
                return None
            n[d] %= self.nc[d]
        return tuple(n)

    def push(self, p):
        s = self.deform * np.prod(np.sin(self.freq * np.pi * (p - self.lo) / (self.hi - self.lo)),
                                  axis=-1)
        return p + s[..., None]

    def map(self, cc, xi):
        """positions [m,3] and Jacobians [m,3,3] at reference points xi [m,3]"""
        cc = np.array(cc)
        if self.mbasis is None:
            x = self.lo + self.h * (cc + xi)
            J = np.broadcast_to(np.diag(self.h), (len(xi), 3, 3)).copy()
            return x, J
        nodes = self.mnodes
        m = len(nodes)
        X = np.zeros((m, m, m, 3))
        for k in range(m):
            for j in range(m):
                for i in range(m):
                    X[i, j, k] = self.push(self.lo + self.h * (cc + np.array([nodes[i], nodes[j], nodes[k]])))
        vx, vy, vz = (self.mbasis.val(xi[:, d]) for d in range(3))
        dx, dy, dz = (self.mbasis.der(xi[:, d]) for d in range(3))
        x = np.einsum("qi,qj,qk,ijkc->qc", vx, vy, vz, X)
        J = np.stack([np.einsum("qi,qj,qk,ijkc->qc", dx, vy, vz, X),
                      np.einsum("qi,qj,qk,ijkc->qc", vx, dy, vz, X),
                      np.einsum("qi,qj,qk,ijkc->qc", vx, vy, dz, X)], axis=-1)
        return x, J


def _cell_basis(b, xi):
    """values [m, n^3] and reference gradients [m, n^3, 3] at points xi [m,3]; x fastest"""
    v = [b.val(xi[:, d]) for d in range(3)]
    g = [b.der(xi[:, d]) for d in range(3)]
    val = np.einsum("qi,qj,qk->qkji", v[0], v[1], v[2]).reshape(len(xi), -1)
    gx = np.einsum("qi,qj,qk->qkji", g[0], v[1], v[2]).reshape(len(xi), -1)
    gy = np.einsum("qi,qj,qk->qkji", v[0], g[1], v[2]).reshape(len(xi), -1)
    gz = np.einsum("qi,qj,qk->qkji", v[0], v[1], g[2]).reshape(len(xi), -1)
    return val, np.stack([gx, gy, gz], axis=-1)


def _face_points(d, s, gp, gw):
    t = [e for e in range(3) if e != d]
    pts, wts = [], []
    for b_ in range(len(gp)):
        for a_ in range(len(gp)):
            xi = np.zeros(3)
            xi[d] = float(s)
            xi[t[0]], xi[t[1]] = gp[a_], gp[b_]
            pts.append(xi)
            wts.append(gw[a_] * gw[b_])
    return np.array(pts), np.array(wts)


def _face_geo(mesh, cc, d, s, gp, gw):
    xi, w = _face_points(d, s, gp, gw)
    x, J = mesh.map(cc, xi)
    Ji = np.linalg.inv(J)
    det = np.linalg.det(J)
    nref = np.zeros(3)
    nref[d] = 1.0 if s else -1.0
    nv = np.einsum("qji,j->qi", Ji, nref)  # J^{-T} nhat
    nrm = np.linalg.norm(nv, axis=1)
    return xi, x, Ji, det * nrm * w, nv / nrm[:, None]


def assemble(mesh, degree, ip=1.0, omega=None, plant=None):
    """dense A (and b for u = prod sin(omega x) if omega is given). `plant` assembles a
    deliberately wrong operator for the tests that show a comparison catches plausible
    mistakes: "drop_symmetry" (no -[u]{dv/dn} term on interior faces), "penalty_min" (min
    instead of max of the cell penalties on interior faces)."""
    n = degree + 1
    n3 = n ** 3
    b1 = Basis1D(gl_nodes(n))
    gp, gw = gauss(n)
    ncell = len(mesh.cells)
    N = ncell * n3
    A = np.zeros((N, N))
    rhs = np.zeros(N) if omega is not None else None
    qx = np.array([[gp[i], gp[j], gp[k]] for k in range(n) for j in range(n) for i in range(n)])
    qw = np.array([gw[i] * gw[j] * gw[k] for k in range(n) for j in range(n) for i in range(n)])
    val, gref = _cell_basis(b1, qx)
    tau = {}
    for cc in mesh.cells:
        x, J = mesh.map(cc, qx)
        Ji = np.linalg.inv(J)
        jxw = np.linalg.det(J) * qw
        gphys = np.einsum("qji,qmj->qmi", Ji, gref)  # J^{-T} grad_ref
        i0 = mesh.index(cc) * n3
        A[i0:i0 + n3, i0:i0 + n3] += np.einsum("q,qmi,qli->ml", jxw, gphys, gphys)
        if omega is not None:
            u = np.prod(np.sin(omega * x), axis=1)
            rhs[i0:i0 + n3] += val.T @ (jxw * 3 * omega ** 2 * u)
        surf = 0.0
        for d in range(3):
            for s in range(2):
                _, _, _, fj, _ = _face_geo(mesh, cc, d, s, gp, gw)
                surf += fj.sum() * (1.0 if mesh.neighbor(cc, d, s) is None else 0.5)
        tau[cc] = surf / jxw.sum()
    pf = ip * (degree + 1) ** 2

    def side(cc, d, s, normal):
        xi, x, Ji, fj, nv = _face_geo(mesh, cc, d, s, gp, gw)
        v, g = _cell_basis(b1, xi)
        dn = np.einsum("qji,qmj,qi->qm", Ji, g, normal)
        return v, dn, x, fj, nv

    for cc in mesh.cells:
        for d in range(3):
            nb = mesh.neighbor(cc, d, 1)
            if nb is not None:  # interior face (cc side 1) -- (nb side 0); normal from cc
                _, _, _, fj, nv = _face_geo(mesh, cc, d, 1, gp, gw)
                vm, dm, _, _, _ = side(cc, d, 1, nv)
                vp, dp, _, _, _ = side(nb, d, 0, nv)
                im, ip_ = mesh.index(cc) * n3, mesh.index(nb) * n3
                jmp = np.concatenate([vm, -vp], axis=1)
                avg = 0.5 * np.concatenate([dm, dp], axis=1)
                t = pf * (min if plant == "penalty_min" else max)(tau[cc], tau[nb])
                sym = 0.0 if plant == "drop_symmetry" else 1.0
                loc = np.einsum("q,qa,qb->ab", fj, -jmp, avg) \
                    + sym * np.einsum("q,qa,qb->ab", fj, -avg, jmp) \
                    + t * np.einsum("q,qa,qb->ab", fj, jmp, jmp)
                # block-wise so that a cell that is its own periodic neighbour accumulates
                A[im:im + n3, im:im + n3] += loc[:n3, :n3]
                A[im:im + n3, ip_:ip_ + n3] += loc[:n3, n3:]
                A[ip_:ip_ + n3, im:im + n3] += loc[n3:, :n3]
                A[ip_:ip_ + n3, ip_:ip_ + n3] += loc[n3:, n3:]
            for s in range(2):
                if mesh.neighbor(cc, d, s) is not None:
                    continue
                _, _, _, fj, nv = _face_geo(mesh, cc, d, s, gp, gw)
                v, dn, x, _, _ = side(cc, d, s, nv)
                i0 = mesh.index(cc) * n3
                t = pf * tau[cc]
                if mesh.bc[2 * d + s] == 0:
                    A[i0:i0 + n3, i0:i0 + n3] += np.einsum("q,qa,qb->ab", fj, -v, dn) \
                        + np.einsum("q,qa,qb->ab", fj, -dn, v) + 2 * t * np.einsum("q,qa,qb->ab", fj, v, v)
                    if omega is not None:
                        g = np.prod(np.sin(omega * x), axis=1)
                        rhs[i0:i0 + n3] += (fj * g) @ (-dn + 2 * t * v)
                elif omega is not None:
                    s3 = np.sin(omega * x)
                    c3 = np.cos(omega * x)
                    grad = omega * np.stack([c3[:, 0] * s3[:, 1] * s3[:, 2], s3[:, 0] * c3[:, 1] * s3[:, 2],
                                             s3[:, 0] * s3[:, 1] * c3[:, 2]], axis=1)
                    h = np.einsum("qi,qi->q", grad, nv)
                    rhs[i0:i0 + n3] += (fj * h) @ v
    return A, rhs


def continuous_nodes(mesh, degree, plant=None):
    """Global node of every (cell, local GL node) of the continuous FE_Q(degree) space on the
    box, x fastest: (n_d degree + 1) nodes per non-periodic direction, n_d degree per periodic
    one; and the Dirichlet mask. Returns (node_of [n_cells, n^3], n_nodes, constrained)."""
    n = degree + 1
    nn = [int(mesh.nc[d]) * degree + (0 if mesh.periodic[d] else 1) for d in range(3)]
    node_of = np.zeros((len(mesh.cells), n ** 3), dtype=np.int64)
    for cc in mesh.cells:
        ci = mesh.index(cc)
        for l, (k, j, i) in enumerate(itertools.product(range(n), range(n), range(n))):
            if plant == "swap_xy":  # a plausible mistake: local node order y fastest
                i, j = j, i
            g = [cc[0] * degree + i, cc[1] * degree + j, cc[2] * degree + k]
            g = [g[d] % nn[d] if mesh.periodic[d] else g[d] for d in range(3)]
            node_of[ci, l] = g[0] + nn[0] * (g[1] + nn[1] * g[2])
    N = nn[0] * nn[1] * nn[2]
    constrained = np.zeros(N, dtype=bool)
    for g in range(N):
        c = (g % nn[0], (g // nn[0]) % nn[1], g // (nn[0] * nn[1]))
        for d in range(3):
            if mesh.periodic[d]:
                continue
            if (c[d] == 0 and mesh.bc[2 * d] == 0) or (c[d] == nn[d] - 1 and mesh.bc[2 * d + 1] == 0):
                constrained[g] = True
    return node_of, N, constrained


def assemble_continuous(mesh, degree, plant=None):
    """dense continuous Q_degree Laplace: cell stiffness matrices (full QGauss(degree+1)
    quadrature) summed through the node map; constrained rows and columns replaced by the
    identity (Dirichlet nodes). plant="swap_xy" assembles with a wrong local node order."""
    n = degree + 1
    b1 = Basis1D(gl_nodes(n))
    gp, gw = gauss(n)
    qx = np.array([[gp[i], gp[j], gp[k]] for k in range(n) for j in range(n) for i in range(n)])
    qw = np.array([gw[i] * gw[j] * gw[k] for k in range(n) for j in range(n) for i in range(n)])
    _, gref = _cell_basis(b1, qx)
    node_of, N, constrained = continuous_nodes(mesh, degree, plant)
    A = np.zeros((N, N))
    for cc in mesh.cells:
        _, J = mesh.map(cc, qx)
        Ji = np.linalg.inv(J)
        jxw = np.linalg.det(J) * qw
        gphys = np.einsum("qji,qmj->qmi", Ji, gref)
        K = np.einsum("q,qmi,qli->ml", jxw, gphys, gphys)
        idx = node_of[mesh.index(cc)]
        np.add.at(A, (idx[:, None], idx[None, :]), K)  # duplicates on a one-cell periodic seam
    A[constrained, :] = 0.0
    A[:, constrained] = 0.0
    A[constrained, constrained] = 1.0
    return A


def assemble_divergence(mesh, degree_p, degree_u):
    """Dense weak DG divergence B (pressure x velocity) of the pressure step
    (divergence_operator.cpp:168-298, weak form, central flux):
      (B u)_i = -sum_K int u . grad q_i + sum_F int {u}.n q_i
    with QGauss(degree_u + 1) on cells and faces. Velocity DoFs per cell: three component blocks
    of (degree_u+1)^3 nodal values. Boundary: where the pressure side is Neumann the velocity is
    Dirichlet at rest (u+ = -u-, no flux); where it is Dirichlet the velocity is Neumann
    (u+ = u-, flux u-.n)."""
    npp, nu = degree_p + 1, degree_u + 1
    n3p, n3u = npp ** 3, nu ** 3
    bp, bu = Basis1D(gl_nodes(npp)), Basis1D(gl_nodes(nu))
    gp, gw = gauss(nu)
    ncell = len(mesh.cells)
    B = np.zeros((ncell * n3p, ncell * 3 * n3u))
    qx = np.array([[gp[i], gp[j], gp[k]] for k in range(nu) for j in range(nu) for i in range(nu)])
    qw = np.array([gw[i] * gw[j] * gw[k] for k in range(nu) for j in range(nu) for i in range(nu)])
    vu, _ = _cell_basis(bu, qx)
    _, gpref = _cell_basis(bp, qx)
    for cc in mesh.cells:
        _, J = mesh.map(cc, qx)
        Ji = np.linalg.inv(J)
        jxw = np.linalg.det(J) * qw
        gphys = np.einsum("qji,qmj->qmi", Ji, gpref)  # grad q_m at the points
        ip, iu = mesh.index(cc) * n3p, mesh.index(cc) * 3 * n3u
        for k in range(3):
            B[ip:ip + n3p, iu + k * n3u:iu + (k + 1) * n3u] -= np.einsum("q,qm,ql->ml", jxw,
                                                                         gphys[:, :, k], vu)
    for cc in mesh.cells:
        for d in range(3):
            for s in range(2):
                nb = mesh.neighbor(cc, d, s)
                xi, _, _, fj, nv = _face_geo(mesh, cc, d, s, gp, gw)
                qm = _cell_basis(bp, xi)[0]  # own pressure test at the face points
                um = _cell_basis(bu, xi)[0]
                ip, iu = mesh.index(cc) * n3p, mesh.index(cc) * 3 * n3u
                if nb is not None:
                    xin = xi.copy()
                    xin[:, d] = 1.0 - s  # the neighbour's side of the shared face
                    up = _cell_basis(bu, xin)[0]
                    jn = mesh.index(nb) * 3 * n3u
                    for k in range(3):
                        B[ip:ip + n3p, iu + k * n3u:iu + (k + 1) * n3u] += \
                            0.5 * np.einsum("q,qm,ql->ml", fj * nv[:, k], qm, um)
                        B[ip:ip + n3p, jn + k * n3u:jn + (k + 1) * n3u] += \
                            0.5 * np.einsum("q,qm,ql->ml", fj * nv[:, k], qm, up)
                elif mesh.bc[2 * d + s] == 0:  # pressure Dirichlet: velocity Neumann
                    for k in range(3):
                        B[ip:ip + n3p, iu + k * n3u:iu + (k + 1) * n3u] += \
                            np.einsum("q,qm,ql->ml", fj * nv[:, k], qm, um)
    return B
