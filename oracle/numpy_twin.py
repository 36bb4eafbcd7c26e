"""Dense-matrix restatement of the projector in world coordinates.

Test infrastructure only. Independent of ``joseph_oracle.c``: it works with
world-space points and planes instead of grid coordinates, and builds the
whole system matrix ``A`` (rows = sinogram ``(i, a, j)``, columns = volume
``(z, y, x)``), so the transpose is ``A.T`` by construction. Only for tiny
problems (``rays * voxels`` up to a few million).

Per ray (pixel centre or detector sub-pixel ``p``, direction ``r``): the
dominant axis ``m`` is chosen per angle from the central direction scaled by
the voxel size (ties x, then y, then z); for every voxel-centre plane
``X_m = min_m + (n + 1/2) h_m`` the intersection ``X`` is bilinearly spread on
the four nearest voxel centres of that plane (zero outside) with weight
``h_m |r| / |r_m|``, the world length of the ray between consecutive planes.
"""
import numpy as np


def _unpack(g):
    V = g.vectors
    return V[:, 0:3], V[:, 3:6], V[:, 6:9], V[:, 9:12]


def system_matrix(g):
    first, det, u, v = _unpack(g)
    n = np.array([g.nx, g.ny, g.nz])
    lo, hi = np.array(g.vmin, float), np.array(g.vmax, float)
    h = (hi - lo) / n
    rows, cols, na, k = g.rows, g.cols, g.n_angles, g.det_supersampling
    A = np.zeros((rows * na * cols, g.nz * g.ny * g.nx))
    ii, jj = np.meshgrid(np.arange(rows), np.arange(cols), indexing="ij")
    ii, jj = ii.ravel(), jj.ravel()
    strides = np.array([1, g.nx, g.nx * g.ny])
    for a in range(na):
        central = (det[a] - first[a]) if g.kind == 1 else first[a]
        e = np.abs(central / h)
        m = 0 if (e[0] >= e[1] and e[0] >= e[2]) else (1 if e[1] >= e[2] else 2)
        c1, c2 = [c for c in (2, 1, 0) if c != m][:2]  # any order works for the weights
        row_index = (ii * na + a) * cols + jj
        for p in range(k):
            for q in range(k):
                fu = jj + (q + 0.5) / k - cols / 2
                fv = ii + (p + 0.5) / k - rows / 2
                P = det[a][None, :] + fu[:, None] * u[a][None, :] + fv[:, None] * v[a][None, :]
                r = (P - first[a][None, :]) if g.kind == 1 else np.broadcast_to(first[a], P.shape)
                ok = r[:, m] != 0
                rm = np.where(ok, r[:, m], 1.0)
                length = h[m] * np.linalg.norm(r, axis=1) / np.abs(rm)
                for sl in range(n[m]):
                    plane = lo[m] + (sl + 0.5) * h[m]
                    t = (plane - P[:, m]) / rm
                    X = P + t[:, None] * r
                    f1 = (X[:, c1] - lo[c1]) / h[c1] - 0.5
                    f2 = (X[:, c2] - lo[c2]) / h[c2] - 0.5
                    b1, b2 = np.floor(f1), np.floor(f2)
                    for d1 in (0, 1):
                        for d2 in (0, 1):
                            k1, k2 = (b1 + d1).astype(np.int64), (b2 + d2).astype(np.int64)
                            w = (1 - np.abs(f1 - k1)) * (1 - np.abs(f2 - k2))
                            good = ok & (k1 >= 0) & (k1 < n[c1]) & (k2 >= 0) & (k2 < n[c2]) & (w > 0)
                            col = sl * strides[m] + k1 * strides[c1] + k2 * strides[c2]
                            np.add.at(A, (row_index[good], col[good]),
                                      (w * length)[good] / (k * k))
    return A


def fp(g, vol):
    return (system_matrix(g) @ np.asarray(vol, float).ravel()).reshape(g.sino_shape)


def bp(g, sino):
    return (system_matrix(g).T @ np.asarray(sino, float).ravel()).reshape(g.vol_shape)
