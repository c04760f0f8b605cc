"""Checks of BEST's transport algebra (best/difco.f:51-401) that the mathematics fixes, for any
result dict with the tensors of bhg_result (the GPU path) or of the oracle's best_run.

None of them restates BEST's formulas; each follows from what the quantities are:

* the mobility at the centre of resistance is the inverse of the 6 x 6 friction matrix there
  [[Ktt, Krt^T], [Krt, Krr]], and the mobility at the centre of diffusion is the same operator
  seen from another point, T M T^T with T = [[I, -[u]x], [0, I]], u = CD - CR (rigid-body
  kinematics: U_P = U_O + w x (P - O), torques move the other way);
* the centre of resistance is the point where the coupling friction is symmetric, the centre of
  diffusion the point where the coupling mobility is symmetric (Brenner 1967); BEST builds the
  3 x 3 systems for them from one triangle of Ktt and Drr, so the symmetry holds to the size of
  those tensors' own asymmetry (the collocation matrix is not exactly symmetric) times |u|;
* lengths: Dtt ~ 1/s, Dtr ~ 1/s^2, Drr ~ 1/s^3, Ktt ~ s, Krt ~ s^2, Krr ~ s^3, centres ~ s,
  CHI V ~ s^3 when every coordinate is scaled by s (without the area correction).
"""
import numpy as np


def cross_matrix(u):
    """[u]x: [u]x v = u x v."""
    return np.array([[0.0, -u[2], u[1]], [u[2], 0.0, -u[0]], [-u[1], u[0], 0.0]])


def _asym_vec(a):
    return np.array([a[2, 1] - a[1, 2], a[0, 2] - a[2, 0], a[1, 0] - a[0, 1]])


def _rel_asym(a):
    return float(np.max(np.abs(a - a.T)) / np.max(np.abs(a)))


def mobility_errors(r):
    """Relative errors of Dtt, Dtr, Drr (at the centre of diffusion) against the inverse of the
    friction matrix at the centre of resistance moved to the centre of diffusion."""
    ktt, krt, krr = (np.asarray(r[k]) for k in ("ktt", "krt", "krr"))
    fric = np.block([[ktt, krt.T], [krt, krr]])
    m_cr = np.linalg.inv(fric)
    t = np.block([[np.eye(3), -cross_matrix(np.asarray(r["udr"]))], [np.zeros((3, 3)), np.eye(3)]])
    m_cd = t @ m_cr @ t.T
    out = {}
    for name, blk, want in (("dtt", m_cd[:3, :3], r["dtt"]), ("dtr", m_cd[:3, 3:], r["dtr"]),
                            ("drr", m_cd[3:, 3:], r["drr"])):
        want = np.asarray(want)
        out[name] = float(np.max(np.abs(blk - want)) / np.max(np.abs(want)))
    # BEST's Schur formulas take the coupling mobility as Dtr^T where the inverse has its own
    # (r, t) block, which differs by the tensors' asymmetry times the coupling; the tt block
    # carries that difference, so its bound is 10 x asymmetry x coupling (+ rounding)
    coupling = float(np.max(np.abs(krt)) / np.sqrt(np.max(np.abs(ktt)) * np.max(np.abs(krr))))
    asym = max(_rel_asym(ktt), _rel_asym(krr), _rel_asym(np.asarray(r["drr"])))
    out["dtt_bound"] = 10.0 * asym * coupling + 1e-13
    return out


def centre_errors(r):
    """(CR error, CR bound, CD error, CD bound): the centres against the points that make the
    coupling tensors symmetric exactly (solved with the full 3 x 3 tensors), and the bound the
    tensors' own asymmetry allows (100 x relative asymmetry x |u| + 1e-12 x |body|)."""
    ktt, krt = np.asarray(r["ktt"]), np.asarray(r["krt"])
    urc = np.asarray(r["urc"])
    k0 = krt - ktt @ cross_matrix(urc)  # the coupling at the centroid, symmetrized as BEST does
    jac = np.column_stack([_asym_vec(ktt @ cross_matrix(e)) for e in np.eye(3)])
    u_cr = np.linalg.solve(jac, -_asym_vec(k0))
    drr, dtr = np.asarray(r["drr"]), np.asarray(r["dtr"])
    udr = np.asarray(r["udr"])
    # Dtr(P) = Dtr(CR) - [P - CR]x Drr: recover Dtr at the centre of resistance, then the point
    # where it becomes symmetric
    d0 = dtr + cross_matrix(udr) @ drr
    jac2 = np.column_stack([_asym_vec(-cross_matrix(e) @ drr) for e in np.eye(3)])
    u_cd = np.linalg.solve(jac2, -_asym_vec(d0))
    size = max(1.0, float(np.max(np.abs(r.get("centroid", np.zeros(3))))))
    b_cr = 100 * _rel_asym(ktt) * max(np.max(np.abs(urc)), 1e-3) + 1e-12 * size
    b_cd = 100 * _rel_asym(drr) * max(np.max(np.abs(udr)), 1e-3) + 1e-12 * size
    return (float(np.max(np.abs(u_cr - urc))), b_cr, float(np.max(np.abs(u_cd - udr))), b_cd)


SCALING = (("dtt", -1), ("dtr", -2), ("drr", -3), ("ktt", 1), ("krt", 2), ("krr", 3), ("urc", 1), ("udr", 1),
           ("dtrans", -1), ("dr", -3), ("chiprime", 3))


def scaling_errors(r1, rs, s):
    """Relative errors of each quantity of rs (the mesh scaled by s) against r1 scaled by s^p."""
    out = {}
    for k, p in SCALING:
        a, b = np.asarray(r1[k], dtype=np.float64), np.asarray(rs[k], dtype=np.float64)
        ref = a * s ** p
        out[k] = float(np.max(np.abs(b - ref)) / max(np.max(np.abs(ref)), 1e-300))
    return out


def cyclic(v):
    """The exact rotation x -> y -> z -> x applied to NAR x 3 vertex arrays: (x, y, z) -> (z, x, y)."""
    return np.asfortranarray(np.asarray(v)[:, [2, 0, 1]])


CYCLIC = np.array([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])


def covariance_errors(r, rr):
    """rr: the result on the cyclically rotated mesh.  Relative errors of R T R^T against its
    tensors and of the scalars."""
    out = {}
    for k in ("dtt", "drr", "ktt", "krr"):
        a = CYCLIC @ np.asarray(r[k]) @ CYCLIC.T
        out[k] = float(np.max(np.abs(a - np.asarray(rr[k]))) / np.max(np.abs(a)))
    for k in ("dtrans", "dr", "chiprime"):
        out[k] = abs(rr[k] - r[k]) / abs(r[k])
    return out


def eigen_errors(r):
    """JACOBI (best/jacobi.f) reads the upper triangle of Dtt and Drr: the eigenvalues against
    LAPACK's of the symmetric matrix with that upper triangle, as sets."""
    out = {}
    for k, e in (("dtt", "dtt_eig"), ("drr", "drr_eig")):
        a = np.asarray(r[k])
        sym = np.triu(a) + np.triu(a, 1).T
        want = np.linalg.eigvalsh(sym)
        got = np.sort(np.asarray(r[e]))
        out[e] = float(np.max(np.abs(got - want)) / np.max(np.abs(want)))
    return out
