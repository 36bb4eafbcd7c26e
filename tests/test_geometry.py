"""Geometry layer: property tests restated from tomosipo's geometry tests."""
import numpy as np
import pytest

import tomosipo_amd as ts
from tomosipo_amd import vector_calc as vc


def test_volume_basics():
    vg = ts.volume(shape=(4, 5, 6), pos=(1, 2, 3), size=(2, 5, 3))
    assert vg.shape == (4, 5, 6)
    assert vg.size == (2.0, 5.0, 3.0)
    assert vg.voxel_size == (0.5, 1.0, 0.5)
    assert vg.extent == ((0.0, 2.0), (-0.5, 4.5), (1.5, 4.5))
    assert ts.volume(shape=3).size == (3.0, 3.0, 3.0)
    assert ts.volume(shape=2, extent=((0, 1), (0, 2), (0, 4))).size == (1.0, 2.0, 4.0)
    with pytest.raises(ValueError):
        ts.volume(shape=2, pos=0, extent=(0, 1))
    with pytest.raises(TypeError):
        ts.volume(shape=(1.5, 2, 3))


def test_volume_slicing_and_astra_roundtrip():
    vg = ts.volume(shape=128)
    assert vg[64, :, :] == vg[64]
    assert vg[64].shape == (1, 128, 128)
    vg = ts.volume(shape=(3, 4, 5), pos=(0.1, 0.2, 0.3), size=(1, 2, 3))
    d = vg.to_astra()
    assert (d["GridSliceCount"], d["GridRowCount"], d["GridColCount"]) == (3, 4, 5)
    assert ts.from_astra(d) == vg


def test_parallel_and_cone_angles():
    assert np.allclose(ts.parallel(angles=4).angles, [0, np.pi / 4, np.pi / 2, 3 * np.pi / 4])
    assert np.allclose(ts.cone(angles=4, cone_angle=1).angles, [0, np.pi / 2, np.pi, 3 * np.pi / 2])
    pg = ts.cone(angles=3, shape=(10, 20), size=(2, 4), cone_angle=0.5)
    assert pg.src_orig_dist == pytest.approx(4.0) and pg.src_det_dist == pytest.approx(4.0)


def test_theta_zero_axes_and_project_point():
    """At angle 0: v along +z, u along +x, source on -y (tomosipo test_cone_vec.py:143-173)."""
    shape, size = (10, 40), (30, 80)
    pg = ts.cone(angles=1, shape=shape, size=size, src_orig_dist=10, src_det_dist=10).to_vec()
    assert np.allclose(pg.det_v, [[3.0, 0, 0]])
    assert np.allclose(pg.det_u, [[0, 0, 2.0]])
    assert np.allclose(pg.src_pos, [[0, -10.0, 0]])
    assert np.allclose(pg.project_point((0, 0, 0)), [[0, 0]])
    assert np.allclose(pg.project_point((3.0, 0, 0)), [[1.0, 0.0]])
    assert np.allclose(pg.project_point((0, 0, 2.0)), [[0.0, 1.0]])
    pg = ts.cone(angles=1, shape=(100, 400), size=size, src_orig_dist=10, src_det_dist=10).to_vec()
    assert np.allclose(pg.project_point((0.3, 0, 0)), [[1.0, 0.0]])
    assert np.allclose(pg.project_point((0, 0, 0.2)), [[0.0, 1.0]])
    par = ts.parallel(angles=1, shape=(10, 40), size=size).to_vec()
    assert np.allclose(par.ray_dir, [[0, -1.0, 0]])
    assert np.allclose(par.project_point((0, 5.0, 2.0)), [[0.0, 1.0]])


def test_geom_2vec_rotation_sense():
    """u and the ray rotate together so that u x v stays the detector normal."""
    t = 0.3
    par = ts.parallel(angles=[t], shape=(2, 2), size=(2, 2)).to_vec()
    assert np.allclose(par.ray_dir, [[0, -np.cos(t), np.sin(t)]])
    assert np.allclose(par.det_u, [[0, np.sin(t), np.cos(t)]])
    assert abs(vc.dot(par.ray_dir, par.det_u)[0]) < 1e-15
    cone = ts.cone(angles=[t], shape=(2, 2), size=(2, 2), src_orig_dist=3, src_det_dist=5).to_vec()
    assert np.allclose(cone.src_pos, [[0, -3 * np.cos(t), 3 * np.sin(t)]])
    assert np.allclose(cone.det_pos, [[0, 2 * np.cos(t), -2 * np.sin(t)]])


def test_src_pos_scaling():
    pg1 = ts.cone(size=np.sqrt(2), src_orig_dist=1, src_det_dist=1).to_vec()
    pg2 = ts.cone(size=np.sqrt(2), src_orig_dist=2, src_det_dist=2).to_vec()
    assert np.all(abs(pg1.src_pos * 2 - pg2.src_pos) < ts.epsilon)


def test_vec_astra_roundtrip(rng):
    for _ in range(20):
        shape = tuple(int(s) for s in rng.integers(1, 100, size=2))
        v1, v2, v3, v4 = (rng.uniform(size=(30, 3)) for _ in range(4))
        pg = ts.cone_vec(shape=shape, src_pos=v1, det_pos=v2, det_v=v3, det_u=v4)
        assert pg == ts.from_astra(pg.to_astra())
        pp = ts.parallel_vec(shape=shape, ray_dir=v1, det_pos=v2, det_v=v3, det_u=v4)
        assert pp == ts.from_astra(pp.to_astra())
    d = ts.cone_vec(shape=(2, 3), src_pos=(1, 2, 3), det_pos=(4, 5, 6), det_v=(7, 8, 9),
                    det_u=(10, 11, 12)).to_astra()
    # ASTRA order: [src, det, u, v], each (x, y, z)
    assert np.allclose(d["Vectors"], [[3, 2, 1, 6, 5, 4, 12, 11, 10, 9, 8, 7]])
    assert d["DetectorRowCount"] == 2 and d["DetectorColCount"] == 3


def test_parametric_astra_roundtrip():
    pg = ts.cone(angles=7, shape=(3, 5), size=(1.5, 2.5), src_orig_dist=3, src_det_dist=7)
    assert ts.from_astra(pg.to_astra()) == pg
    pp = ts.parallel(angles=7, shape=(3, 5), size=(1.5, 2.5))
    assert ts.from_astra(pp.to_astra()) == pp


def test_det_sizes_reshape_slicing():
    pg = ts.cone(angles=5, size=(1, 1), cone_angle=0.5).to_vec()
    assert np.allclose(pg.det_sizes, 1.0)
    pg2 = pg.reshape((7, 9))
    assert pg2.det_shape == (7, 9) and np.allclose(pg2.det_sizes, pg.det_sizes)
    assert pg.corners.shape == (5, 4, 3)
    assert ts.cone(angles=10, cone_angle=0.5).to_vec()[0::2].num_angles == 5
    assert ts.cone(shape=10, cone_angle=0.5).to_vec()[:, ::2, ::2].det_shape == (5, 5)
    assert ts.parallel(angles=10)[0::2].num_angles == 5


def test_transform_group_laws(rng):
    T = ts.translate(rng.normal(size=3))
    R = ts.rotate(pos=rng.normal(size=3), axis=rng.normal(size=3), angles=rng.normal())
    S = ts.scale(abs(rng.normal(size=3)) + 0.1, pos=rng.normal(size=3))
    for M in (T, R, S, T * R * S):
        assert M * M.inv == ts.geometry.identity()
    assert (T * R) * S == T * (R * S)
    # right-handed rotation about +z maps +x onto +y
    Rz = ts.rotate(pos=0, axis=(1, 0, 0), angles=np.pi / 2)
    assert np.allclose(Rz.transform_vec((0, 0, 1)), [[0, 1, 0]])
    assert np.allclose(ts.rotate(pos=0, axis=(1, 0, 0), angles=np.pi / 2, right_handed=False)
                       .transform_vec((0, 0, 1)), [[0, -1, 0]])
    # rotations preserve lengths and translate moves points but not vectors
    v = rng.normal(size=(5, 3))
    assert np.allclose(vc.norm(R.transform_vec(v)), vc.norm(v))
    assert np.allclose(T.transform_vec(v), v)
    assert np.allclose(ts.translate((1, 2, 3)).transform_point((0, 0, 0)), [[1, 2, 3]])
    # multi-step broadcasting
    T2 = ts.translate((1, 0, 0), alpha=[0.0, 1.0])
    assert (ts.translate((1, 0, 0)) * T2).num_steps == 2
    with pytest.raises(ValueError):
        T2 * ts.translate((1, 0, 0), alpha=[0.0, 1.0, 2.0])


def test_scale_doctest_matrix():
    M = ts.scale((2, 2, 2), pos=(3, 4, 5)).matrix[0]
    assert np.allclose(M, [[2, 0, 0, -3], [0, 2, 0, -4], [0, 0, 2, -5], [0, 0, 0, 1]])


def test_perspective():
    P = ts.to_perspective(pos=(1, 2, 3), w=(1, 0, 0), v=(0, 1, 0), u=(0, 0, 1))
    assert P == ts.translate((1, 2, 3))
    assert ts.from_perspective(vol=ts.volume(pos=(1, 2, 3))) == ts.translate((1, 2, 3)).inv


def test_reflect():
    Rf = ts.reflect(pos=0, axis=(1, 0, 0))
    assert np.allclose(Rf.transform_point((2, 3, 4)), [[-2, 3, 4]])
    assert Rf * Rf == ts.geometry.identity()


def test_volume_transforms():
    vg = ts.volume(shape=4, pos=(1, 2, 3), size=2)
    assert (ts.translate((1, 1, 1)) * vg) == ts.volume(shape=4, pos=(2, 3, 4), size=2)
    with pytest.warns(UserWarning):
        vv = ts.rotate(pos=0, axis=(1, 0, 0), angles=0.5) * vg
    assert isinstance(vv, ts.geometry.VolumeVectorGeometry)
    assert np.allclose(vv.sizes, [[2, 2, 2]])


def test_concatenate():
    pg = ts.concatenate([ts.cone(angles=3, cone_angle=1), ts.cone(angles=2, cone_angle=1)])
    assert pg.num_angles == 5 and pg.is_cone and pg.is_vec
    pp = ts.concatenate([ts.parallel(angles=3), ts.parallel(angles=3).to_vec()])
    assert pp.num_angles == 6 and pp.is_parallel
    T = ts.concatenate([ts.translate((1, 0, 0)), ts.translate((2, 0, 0))])
    assert T.num_steps == 2
    with pytest.raises(ValueError):
        ts.concatenate([])


def test_phantom_hollow_box_mean():
    """tomosipo tests/test_phantom.py:11-17: mean 0.208 at 100^3."""
    vd = ts.data(ts.volume(shape=100))
    ts.phantom.hollow_box(vd)
    assert vd.data.mean() == pytest.approx(0.208)


def test_astra_compatible_normalisation(rng):
    """A rotated/translated volume_vec becomes an axis-aligned volume on the origin,
    and the same transform is applied to the projection geometry."""
    from tomosipo_amd.Operator import to_astra_compatible_operator_geometry

    vg = ts.volume(shape=(8, 9, 10), size=(1, 1.2, 1.4))
    pg = ts.parallel(angles=5, shape=(9, 13), size=1.5).to_vec()
    T = ts.translate(rng.normal(size=3)) * ts.rotate(pos=0, axis=rng.normal(size=3), angles=0.7)
    vg2, pg2 = to_astra_compatible_operator_geometry(T * vg.to_vec(), pg)
    assert vg2 == ts.volume(shape=vg.shape, size=vg.size)
    assert np.allclose(pg2.ray_dir, T.inv.transform_vec(pg.ray_dir))
    assert np.allclose(pg2.det_pos, T.inv.transform_point(pg.det_pos))
    with pytest.raises(TypeError):
        to_astra_compatible_operator_geometry("not a volume", pg)
