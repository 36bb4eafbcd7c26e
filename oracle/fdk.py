"""FP64 restatement of the FDK reconstruction (test infrastructure only).

tomosipo reaches FDK through ``astra.experimental.accumulate_FDK``
(tomosipo/astra.py:374-406), i.e. ASTRA's CUDA FDK, which is not available
here; its bit-level output is therefore parity unpinned. This module states
the Feldkamp-Davis-Kress algorithm (Feldkamp, Davis & Kress, J. Opt. Soc. Am.
A 1(6), 1984; discretisation as in Kak & Slaney, *Principles of Computerized
Tomographic Imaging*, ch. 3) for cone_vec geometries, as the library
implements it, in world coordinates:

* per angle a: source s; unit detector normal n (pointing away from the
  source); the detector is moved to the parallel plane through the volume
  centre c (virtual detector: pixel (0, 0) centre P0, column step U, row
  step V); D = (c - s) . n;
* pre-weighting: y'[i, a, j] = y[i, a, j] * D / |P0 + j U + i V - s|;
* ramp filter along each detector row, linear (zero-padded) convolution with
  the sampled Ram-Lak kernel g[0] = 1/4, g[odd m] = -1/(pi^2 m^2), g[even] = 0,
  scaled by 1/|U|: q[j] = (1/|U|) sum_k g[j - k] y'[k];
* backprojection: f(X) = sum_a (dbeta_a / 2) * lambda^2 * q_a(i*, j*), where
  lambda = D / ((X - s) . n), (j*, i*) are the virtual-detector coordinates
  of the ray s -> X, q is interpolated bilinearly (zero outside) and dbeta_a
  is the source's angular step around the z axis through c (central
  differences of the unwrapped azimuth; one-sided at the ends).

For a full circular scan this is the textbook FDK; as the cone angle goes
to 0 it becomes filtered backprojection. ``tests/test_fdk.py`` pins the
normalisation with a reconstruction of a uniform ball (interior value 1) and
the filter with the known response of the Ram-Lak kernel to a constant row.
"""
import numpy as np


def ramp_kernel(n):
    """Sampled Ram-Lak kernel g[m] for lags m = -(n-1) .. n-1 (unit spacing)."""
    m = np.arange(-(n - 1), n)
    g = np.zeros(m.shape, dtype=np.float64)
    g[m == 0] = 0.25
    odd = (m % 2) != 0
    g[odd] = -1.0 / (np.pi ** 2 * m[odd].astype(np.float64) ** 2)
    return g


def angle_steps(src, centre):
    """Angular step of the source around the z axis through ``centre``."""
    n = len(src)
    if n < 2:
        return np.full(n, 2 * np.pi)
    beta = np.unwrap(np.arctan2(src[:, 1] - centre[1], src[:, 0] - centre[0]))
    d = np.empty(n)
    d[1:-1] = np.abs(beta[2:] - beta[:-2]) / 2
    d[0] = abs(beta[1] - beta[0])
    d[-1] = abs(beta[-1] - beta[-2])
    return d


def params(g):
    """Per-angle FDK quantities in world (x, y, z) coordinates."""
    if g.kind != 1:
        raise ValueError("FDK needs a cone-beam geometry")
    V = g.vectors
    lo, hi = np.array(g.vmin, float), np.array(g.vmax, float)
    c = 0.5 * (lo + hi)
    out = []
    dbeta = angle_steps(V[:, 0:3], c)
    for a in range(g.n_angles):
        s, det, u, v = V[a, 0:3], V[a, 3:6], V[a, 6:9], V[a, 9:12]
        nrm = np.cross(u, v)
        nrm /= np.linalg.norm(nrm)
        if np.dot(det - s, nrm) < 0:
            nrm = -nrm
        P0 = det + (0.5 - g.cols / 2) * u + (0.5 - g.rows / 2) * v
        lam = np.dot(c - s, nrm) / np.dot(P0 - s, nrm)
        P0v, Uv, Vv = s + lam * (P0 - s), lam * u, lam * v
        D = np.dot(c - s, nrm)
        VxN, UxN = np.cross(Vv, nrm), np.cross(Uv, nrm)
        out.append(dict(s=s, n=nrm, D=D, P0=P0v, U=Uv, V=Vv, Bj=VxN / np.dot(VxN, Uv),
                        Bi=UxN / np.dot(UxN, Vv), half_dbeta=dbeta[a] / 2))
    return out


def filtered(g, sino, prm=None):
    """Cosine-weighted, ramp-filtered projections q[i, a, j] (FP64)."""
    prm = params(g) if prm is None else prm
    sino = np.asarray(sino, dtype=np.float64)
    rows, cols = g.rows, g.cols
    kern = ramp_kernel(cols)
    q = np.empty_like(sino)
    ii, jj = np.meshgrid(np.arange(rows), np.arange(cols), indexing="ij")
    for a, p in enumerate(prm):
        P = p["P0"][None, None, :] + jj[..., None] * p["U"] + ii[..., None] * p["V"]
        cw = p["D"] / np.linalg.norm(P - p["s"], axis=-1)
        yw = sino[:, a, :] * cw
        inv_du = 1.0 / np.linalg.norm(p["U"])
        for i in range(rows):
            # full linear convolution; output index j sits at lag offset cols-1
            q[i, a, :] = inv_du * np.convolve(yw[i], kern)[cols - 1:2 * cols - 1]
    return q


def fdk(g, sino, out=None):
    """FDK reconstruction of ``sino`` ([row][angle][col]) into a [z][y][x] volume."""
    prm = params(g)
    q = filtered(g, sino, prm)
    lo, hi = np.array(g.vmin, float), np.array(g.vmax, float)
    n = np.array([g.nx, g.ny, g.nz])
    h = (hi - lo) / n
    zz, yy, xx = np.meshgrid(np.arange(g.nz), np.arange(g.ny), np.arange(g.nx), indexing="ij")
    X = np.stack([lo[0] + (xx + 0.5) * h[0], lo[1] + (yy + 0.5) * h[1], lo[2] + (zz + 0.5) * h[2]], -1)
    f = np.zeros(g.vol_shape)
    for a, p in enumerate(prm):
        W = X - p["s"]
        lam = p["D"] / (W @ p["n"])
        H = p["s"] + lam[..., None] * W - p["P0"]
        j = H @ p["Bj"]
        i = H @ p["Bi"]
        j0, i0 = np.floor(j).astype(np.int64), np.floor(i).astype(np.int64)
        wj, wi = j - j0, i - i0
        val = np.zeros(g.vol_shape)
        qa = q[:, a, :]
        for di, wgt_i in ((0, 1 - wi), (1, wi)):
            for dj, wgt_j in ((0, 1 - wj), (1, wj)):
                ri, cj = i0 + di, j0 + dj
                ok = (ri >= 0) & (ri < g.rows) & (cj >= 0) & (cj < g.cols)
                val += np.where(ok, qa[np.clip(ri, 0, g.rows - 1), np.clip(cj, 0, g.cols - 1)], 0.0) \
                    * wgt_i * wgt_j
        f += p["half_dbeta"] * lam ** 2 * val
    if out is not None:
        out += f
        return out
    return f
