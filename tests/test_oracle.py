"""Pinning the FP64 oracle (CPU).

Each check is chosen so that a plausible convention mistake fails it:
closed forms (axis-aligned sums, chord lengths, a single voxel), an
independent world-coordinate dense-matrix restatement, the exact adjoint, and
the transform equivariance that tomosipo's own operator tests assert
(tests/test_Operator.py:52-112).
"""
import numpy as np
import pytest

import oracle
from oracle import numpy_twin
import tomosipo_amd as ts


def _small_parallel():
    vg = ts.volume(shape=64, size=1)
    pg = ts.parallel(angles=96, shape=(64, 96), size=(1, 1.5))
    return vg, pg


def test_closed_form_theta_zero_and_ninety():
    """Parallel beam, pixels aligned with voxels: y[i, a, j] = h * sum along the ray axis."""
    vg, pg = _small_parallel()
    g = oracle.oracle_geometry(vg, pg)
    x = np.random.default_rng(0).random(g.vol_shape)
    y = oracle.fp(g, x)
    # angle 0: rays along -y, u along +x; pixel j covers voxel column j - 16
    np.testing.assert_allclose(y[:, 0, 16:80], x.sum(axis=1) / 64, rtol=1e-13, atol=1e-15)
    assert np.all(y[:, 0, :16] == 0) and np.all(y[:, 0, 80:] == 0)
    # angle pi/2 (index 48): rays along +x, u along +y
    np.testing.assert_allclose(y[:, 48, 16:80], x.sum(axis=2) / 64, rtol=1e-12, atol=1e-15)


def test_closed_form_chord_length_all_ones():
    """All-ones volume: a ray whose every sample has full support gets n_slices * l."""
    vg = ts.volume(shape=32, size=1)
    t = 0.3
    pg = ts.parallel(angles=[t], shape=(32, 48), size=(1, 1.5))
    g = oracle.oracle_geometry(vg, pg)
    y = oracle.fp(g, np.ones(g.vol_shape))
    # stepping along y (|cos t| > |sin t|); the central pixel ray passes the centre
    length = (1 / 32) * np.sqrt(1 + np.tan(t) ** 2)
    assert y[16, 0, 24] == pytest.approx(32 * length, rel=1e-12)


def test_single_voxel_single_pixel():
    """1-voxel volume, 1-pixel detector through its centre: FP = voxel size."""
    vg = ts.volume(shape=1, size=0.5)
    for t, pg in ((0.2, ts.parallel(angles=[0.2], shape=1, size=0.1)),
                  (1.1, ts.cone(angles=[1.1], shape=1, size=0.1, src_orig_dist=3, src_det_dist=5))):
        g = oracle.oracle_geometry(vg, pg)
        y = oracle.fp(g, np.ones((1, 1, 1)))
        # the ray through the centre samples the only slice at the voxel centre;
        # the stepping axis is the larger of |cos t| (y) and |sin t| (x)
        e = max(abs(np.cos(t)), abs(np.sin(t)))
        assert y[0, 0, 0] == pytest.approx(0.5 / e, rel=1e-12)


def _tiny_geometries(rng):
    out = []
    for k in range(6):
        shape = tuple(int(s) for s in rng.integers(2, 6, size=3))
        vg = ts.volume(shape=shape, pos=rng.normal(0, 0.1, 3), size=rng.uniform(0.5, 1.5, 3))
        T = ts.rotate(pos=0, axis=rng.normal(size=3), angles=rng.normal())
        rows, cols = (int(s) for s in rng.integers(2, 6, size=2))
        angles = rng.uniform(0, 2 * np.pi, 3)
        if k % 2 == 0:
            pg = T * ts.cone(angles=angles, shape=(rows, cols), size=rng.uniform(1, 3, 2),
                             src_orig_dist=rng.uniform(2, 4), src_det_dist=rng.uniform(4, 7)).to_vec()
        else:
            pg = T * ts.parallel(angles=angles, shape=(rows, cols), size=rng.uniform(1, 3, 2)).to_vec()
        out.append((vg, pg, 1 + k % 3))
    return out


def test_numpy_twin_matches_c_oracle(rng):
    for vg, pg, ss in _tiny_geometries(rng):
        g = oracle.oracle_geometry(vg, pg, det_supersampling=ss)
        x = rng.random(g.vol_shape)
        y = rng.random(g.sino_shape)
        np.testing.assert_allclose(oracle.fp(g, x), numpy_twin.fp(g, x), rtol=1e-11, atol=1e-13)
        np.testing.assert_allclose(oracle.bp(g, y), numpy_twin.bp(g, y), rtol=1e-11, atol=1e-13)


def test_adjoint_exact_fp64(rng):
    for vg, pg in [_small_parallel(), (ts.volume(shape=(20, 24, 28), size=(1, 1.1, 1.2)),
                          ts.cone(angles=30, shape=(24, 36), size=(1.6, 2.4), cone_angle=0.5))]:
        g = oracle.oracle_geometry(vg, pg)
        x = rng.random(g.vol_shape)
        y = rng.random(g.sino_shape)
        Ax, ATy = oracle.fp(g, x), oracle.bp(g, y)
        err = abs(np.vdot(Ax, y) - np.vdot(x, ATy)) / (np.linalg.norm(Ax) * np.linalg.norm(y))
        assert err < 1e-13


def test_additive_mode(rng):
    vg, pg = _small_parallel()
    g = oracle.oracle_geometry(vg, pg[::8])
    x = rng.random(g.vol_shape)
    y0 = rng.random(g.sino_shape)
    np.testing.assert_allclose(oracle.fp(g, x, out=y0.copy()), y0 + oracle.fp(g, x), rtol=1e-13)


@pytest.mark.parametrize("seed", [0, 1])
def test_transform_equivariance(seed):
    """op(T vg, pg) == op(vg, T^-1 pg) (tomosipo tests/test_Operator.py:52-83)."""
    from tomosipo_amd.Operator import to_astra_compatible_operator_geometry

    rng = np.random.default_rng(seed)
    vg = ts.volume(shape=12, size=1)
    pg = ts.parallel(angles=16, shape=18, size=1.5).to_vec()
    T = ts.translate(rng.normal(0, 0.1, 3)) * ts.rotate(pos=0, axis=rng.normal(size=3),
                                                        angles=rng.normal())
    vg_a, pg_a = to_astra_compatible_operator_geometry(T * vg.to_vec(), pg)
    vg_b, pg_b = vg, T.inv * pg
    x = rng.random(vg.shape)
    ya = oracle.fp(oracle.oracle_geometry(vg_a, pg_a), x)
    yb = oracle.fp(oracle.oracle_geometry(vg_b, pg_b), x)
    np.testing.assert_allclose(ya, yb, rtol=1e-9, atol=1e-12)


def test_dominant_axis_rule():
    """Parallel beam over [0, pi): y for |cos| >= |sin| (ties x first), x otherwise."""
    vg, pg = _small_parallel()
    axes = oracle.axes(oracle.oracle_geometry(vg, pg))
    t = pg.angles
    ray = np.stack([np.sin(t), -np.cos(t)], axis=1)  # (x, y) components
    expect = np.where(np.abs(ray[:, 0]) >= np.abs(ray[:, 1]), 0, 1)
    np.testing.assert_array_equal(axes, expect)


def test_oracle_catches_planted_sign_error():
    """A flipped detector u (a plausible sign mistake) breaks the theta=0 closed form."""
    vg, pg = _small_parallel()
    bad = ts.parallel_vec(shape=pg.det_shape, ray_dir=pg.ray_dir, det_pos=pg.det_pos,
                          det_v=pg.det_v, det_u=-pg.det_u)
    x = np.random.default_rng(1).random((64, 64, 64))
    y = oracle.fp(oracle.oracle_geometry(vg, bad), x)
    assert not np.allclose(y[:, 0, 16:80], x.sum(axis=1) / 64)


# ---------------------------------------------------------------------------
# cone beam: closed forms that fix the source position, the magnification
# and the detector orientation, with planted mistakes they must catch
# ---------------------------------------------------------------------------
def _cone_geometry(theta, sod=3.0, sdd=5.0, rows=33, cols=49, size=(1.65, 2.45)):
    return ts.cone(angles=[theta], shape=(rows, cols), size=size, src_orig_dist=sod,
                   src_det_dist=sdd).to_vec()


def _cone_chords(pg, vg, pixels):
    """FP of an all-ones unit-cube volume for the given pixels (i, j)."""
    g = oracle.oracle_geometry(vg, pg)
    y = oracle.fp(g, np.ones(g.vol_shape))
    return np.array([y[i, 0, j] for i, j in pixels])


def _chord_closed_form(pg, vg, pixels):
    """Chord n_m * l of the ray from the source through pixel (i, j): every one
    of the n_m slices of the stepping axis m is crossed with full bilinear
    support (the rays stay well inside the cube), l = |h * D| / |D_m|."""
    rows, cols = pg.det_shape
    s, d, u, v = pg.src_pos[0], pg.det_pos[0], pg.det_u[0], pg.det_v[0]
    h = np.array(vg.voxel_size)
    out = []
    for i, j in pixels:
        p = d + (j + 0.5 - cols / 2) * u + (i + 0.5 - rows / 2) * v
        D = (p - s) / h
        m = int(np.argmax(np.abs(D)))
        out.append(vg.shape[m] * np.linalg.norm(h * D) / abs(D[m]))
    return np.array(out)


# the central pixel and pixels up to 4 rows / columns from it: their rays
# cross every y-slice of the cube with full bilinear support at both angles
PIXELS = [(16, 24), (16, 28), (12, 24), (20, 20), (13, 27)]


@pytest.mark.parametrize("theta", [0.0, 0.4])
def test_cone_chord_lengths_all_ones(theta):
    """All-ones unit cube 32^3: FP of a ray equals its chord n_m * l, which for
    the central ray is 1 / cos(theta) and off-centre grows with the ray's
    angle to the central ray (set by the source-detector distance)."""
    vg = ts.volume(shape=32, size=1)
    pg = _cone_geometry(theta)
    got = _cone_chords(pg, vg, PIXELS)
    np.testing.assert_allclose(got, _chord_closed_form(pg, vg, PIXELS), rtol=1e-12)
    assert got[0] == pytest.approx(1 / np.cos(theta), rel=1e-12)  # the central ray
    # a planted source distance 10 % off changes the off-centre chords
    for sod in (2.7, 3.3):
        bad = _cone_geometry(theta, sod=sod, sdd=5.0 + sod - 3.0)  # detector kept in place
        assert not np.allclose(_cone_chords(bad, vg, PIXELS)[1:], got[1:], rtol=1e-7)


def _footprint_centroid(y):
    """Weighted centroid (i, j) of a single-angle sinogram y[i, 0, j]."""
    w = y[:, 0, :]
    ii, jj = np.meshgrid(np.arange(w.shape[0]), np.arange(w.shape[1]), indexing="ij")
    return (w * ii).sum() / w.sum(), (w * jj).sum() / w.sum()


def _magnified_position(pg, X):
    """Detector pixel coordinates (i, j) of s + lam (X - s), the central
    projection of X from the source onto the detector plane."""
    rows, cols = pg.det_shape
    s, d, u, v = pg.src_pos[0], pg.det_pos[0], pg.det_u[0], pg.det_v[0]
    n = np.cross(v, u)
    lam = np.dot(d - s, n) / np.dot(X - s, n)
    p = s + lam * (X - s)
    j = np.dot(p - d, u) / np.dot(u, u) + cols / 2 - 0.5
    i = np.dot(p - d, v) / np.dot(v, v) + rows / 2 - 0.5
    return i, j


@pytest.mark.parametrize("theta", [0.0, 0.7])
def test_cone_off_centre_voxel_projects_to_its_magnified_position(theta):
    """One non-zero voxel off the centre (64^3 unit cube, detector pixels a
    quarter of the magnified voxel): its footprint's centroid is the central
    projection of the voxel centre, i = v-coordinate and j = u-coordinate of
    s + lam (X - s), to 0.05 pixel: the footprint is a tent of about 1.6
    pixels' half-width centred there, and the centroid of its samples at the
    pixel centres is off by a few hundredths of a pixel, depending on where
    the centre falls between pixels. A detector with flipped u, or a source
    10 % nearer or farther, moves it by pixels."""
    vg = ts.volume(shape=64, size=1)
    k = (20, 45, 13)                                  # (z, y, x) voxel index
    X = (np.array(k) + 0.5) / 64 - 0.5                 # its centre (z, y, x)
    x = np.zeros((64, 64, 64))
    x[k] = 1.0
    pg = _cone_geometry(theta, sod=3.0, sdd=6.0, rows=97, cols=97, size=(1.9, 1.9))
    y = oracle.fp(oracle.oracle_geometry(vg, pg), x)
    ci, cj = _footprint_centroid(y)
    ei, ej = _magnified_position(pg, X)
    assert abs(ci - ei) < 0.05 and abs(cj - ej) < 0.05, (ci, ei, cj, ej)
    bads = [ts.cone_vec(shape=pg.det_shape, src_pos=pg.src_pos, det_pos=pg.det_pos, det_v=pg.det_v,
                        det_u=-pg.det_u)]
    for sod in (2.7, 3.3):
        bads.append(_cone_geometry(theta, sod=sod, sdd=6.0 + sod - 3.0, rows=97, cols=97, size=(1.9, 1.9)))
    for bad in bads:
        bi, bj = _footprint_centroid(oracle.fp(oracle.oracle_geometry(vg, bad), x))
        assert max(abs(bi - ei), abs(bj - ej)) > 0.5, (bi, bj, ei, ej)


def test_cone_magnification_closed_form():
    """The footprint centroids of voxels along the source axis scale with the
    magnification SDD / (SOD + depth) (theta = 0: the source at -SOD along y),
    to 0.05 pixel as above."""
    vg = ts.volume(shape=64, size=1)
    pg = _cone_geometry(0.0, sod=3.0, sdd=6.0, rows=97, cols=97, size=(1.9, 1.9))
    pitch = 1.9 / 97
    for ky in (5, 32, 58):
        k = (40, ky, 50)
        X = (np.array(k) + 0.5) / 64 - 0.5
        x = np.zeros((64, 64, 64))
        x[k] = 1.0
        ci, cj = _footprint_centroid(oracle.fp(oracle.oracle_geometry(vg, pg), x))
        M = 6.0 / (3.0 + X[1])                     # depth along +y from the source at y = -3
        assert (ci - 48) * pitch == pytest.approx(M * X[0], abs=0.05 * pitch)
        assert (cj - 48) * pitch == pytest.approx(M * X[2], abs=0.05 * pitch)
