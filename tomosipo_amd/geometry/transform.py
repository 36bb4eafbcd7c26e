"""Homogeneous 4x4 transforms acting on ``(z, y, x)`` coordinates.

Semantics follow tomosipo ``geometry/transform.py``:

* ``Transform`` holds ``(num_steps, 4, 4)`` matrices; products broadcast a
  single step against many (``transform.py:17-54``).
* points transform with homogeneous coordinate 1, vectors with 0
  (``transform.py:79-113``).
* ``translate``/``scale``/``rotate``/``reflect`` (``transform.py:126-452``);
  ``rotate`` is right-handed in the physical ``(x, y, z)`` frame, i.e. a
  rotation about ``+z`` by +90 degrees maps ``+x`` onto ``+y``.
* ``to_perspective`` has matrix columns ``(w, v, u, pos)`` and
  ``from_perspective`` is its inverse (``transform.py:455-661``).
"""
import numpy as np

from .. import vector_calc as vc
from .. import types as tt

_EPS = 1e-8


class Transform:
    def __init__(self, matrix):
        self.matrix = vc.to_matrices(matrix)

    def __mul__(self, other):
        if isinstance(other, Transform):
            a, b = self.num_steps, other.num_steps
            if not (a == 1 or b == 1 or a == b):
                raise ValueError(
                    "Cannot multiply transforms with different number of time steps. "
                    f"Got steps: {a} and {b}"
                )
            return Transform(vc.matrix_matrix_transform(self.matrix, other.matrix))
        return NotImplemented

    def __repr__(self):
        return f"Transform(\n    {self.matrix}\n)"

    def __eq__(self, other):
        if not isinstance(other, Transform):
            return False
        n = max(self.num_steps, other.num_steps)
        try:
            A = np.broadcast_to(self.matrix, (n, 4, 4))
            B = np.broadcast_to(other.matrix, (n, 4, 4))
        except ValueError:
            return False
        return bool(np.all(abs(A - B) < _EPS))

    def __getitem__(self, i):
        if not isinstance(i, (slice, int, np.integer)):
            raise TypeError(f"Transform only support one-dimensional indexing. Got: {i}")
        return Transform(self.matrix[i])

    def __len__(self):
        return self.num_steps

    @property
    def num_steps(self):
        return self.matrix.shape[0]

    @property
    def inv(self):
        return Transform(vc.invert_transformation_matrix(self.matrix))

    def transform_vec(self, vec):
        return vc.from_homogeneous(vc.matrix_transform(self.matrix, tt.to_homogeneous_vec(vec)))

    def transform_point(self, points):
        return vc.from_homogeneous(vc.matrix_transform(self.matrix, tt.to_homogeneous_pos(points)))


def identity():
    return Transform(np.eye(4))


def translate(axis, *, alpha=1):
    axis = tt.to_homogeneous_pos(axis)
    alpha = tt.to_scalars(alpha, var_name="alpha")
    if not (len(axis) == 1 or len(alpha) == 1 or len(axis) == len(alpha)):
        raise ValueError(
            "Expected parameters `axis` and `alpha` to be the same length. "
            f"Got: {len(axis)}, {len(alpha)}"
        )
    t = alpha[:, None] * axis[:, :3]
    M = np.broadcast_to(np.eye(4), (len(t), 4, 4)).copy()
    M[:, :3, 3] = t
    return Transform(M)


def scale(s, *, pos=0, alpha=1.0):
    if np.isscalar(s):
        s = tt.to_size3d(s)
    if np.isscalar(pos):
        pos = tt.to_pos(pos)
    s = tt.to_homogeneous_vec(s)
    pos = tt.to_homogeneous_pos(pos)
    alpha = tt.to_scalars(alpha)
    try:
        n = vc.broadcast_lengths(vc.broadcast_lengths(len(s), len(pos)), len(alpha))
    except ValueError:
        raise ValueError(
            "Expected `scale`, `pos`, and `alpha` to be broadcastable. "
            f"Got lengths: {len(s)}, {len(pos)}, and {len(alpha)}."
        )
    sv = np.broadcast_to(alpha[:, None] * s[:, :3], (n, 3))
    M = np.zeros((n, 4, 4))
    M[:, 0, 0], M[:, 1, 1], M[:, 2, 2], M[:, 3, 3] = sv[:, 0], sv[:, 1], sv[:, 2], 1.0
    T = translate(pos)
    return T * Transform(M) * T.inv


def rotate(*, pos, axis, angles=None, rad=None, deg=None, right_handed=True):
    if np.isscalar(pos):
        pos = tt.to_pos(pos)
    pos = tt.to_homogeneous_pos(pos)
    axis = tt.to_homogeneous_vec(axis)
    n = vc.broadcast_lengths(len(pos), len(axis))
    pos, axis = np.broadcast_to(pos, (n, 4)), np.broadcast_to(axis, (n, 4))
    a = axis[:, :3] / vc.norm(axis[:, :3])[:, None]
    if angles is None and rad is None and deg is None:
        raise ValueError("The `angles=` parameter is required.")
    if angles is not None and (rad is not None or deg is not None):
        raise TypeError(
            "The `angles` parameter is not compatible with the `rad` or `deg` parameter."
        )
    if angles is None:
        angles = np.deg2rad(deg) if deg is not None else rad
    theta = tt.to_scalars(angles, var_name="angles")
    if not right_handed:
        theta = -theta
    m = vc.broadcast_lengths(len(a), len(theta))
    a = np.broadcast_to(a, (m, 3))
    theta = np.broadcast_to(theta, (m,))
    # Rodrigues' formula evaluated in (z, y, x) component order. Because that
    # order is a reflection of (x, y, z), the physically right-handed rotation
    # by theta is the (z, y, x)-Rodrigues rotation by -theta.
    c, s = np.cos(-theta), np.sin(-theta)
    K = np.zeros((m, 3, 3))
    K[:, 0, 1], K[:, 0, 2] = -a[:, 2], a[:, 1]
    K[:, 1, 0], K[:, 1, 2] = a[:, 2], -a[:, 0]
    K[:, 2, 0], K[:, 2, 1] = -a[:, 1], a[:, 0]
    outer = a[:, :, None] * a[:, None, :]
    R3 = c[:, None, None] * np.eye(3) + s[:, None, None] * K + (1 - c)[:, None, None] * outer
    R = np.zeros((m, 4, 4))
    R[:, :3, :3] = R3
    R[:, 3, 3] = 1.0
    T = translate(-pos[:, :3])
    return T.inv * Transform(R) * T


def reflect(*, pos, axis):
    if np.isscalar(pos):
        pos = tt.to_pos(pos)
    pos = tt.to_homogeneous_pos(pos)
    axis = tt.to_homogeneous_vec(axis)
    n = vc.broadcast_lengths(len(pos), len(axis))
    pos, axis = np.broadcast_to(pos, (n, 4)), np.broadcast_to(axis, (n, 4))
    a = axis[:, :3] / vc.norm(axis[:, :3])[:, None]
    M = np.broadcast_to(np.eye(4), (n, 4, 4)).copy()
    M[:, :3, :3] -= 2 * a[:, :, None] * a[:, None, :]
    T = translate(pos[:, :3])
    return T * Transform(M) * T.inv


def to_perspective(*, pos=None, w=None, v=None, u=None, vol=None, ignore_scale=True):
    if vol is not None:
        pos, w, v, u = vol.pos, vol.w, vol.v, vol.u
    if any(x is None for x in (pos, w, v, u)):
        raise ValueError("Not enough arguments provided: one of pos, w, v, u is missing.")
    pos = tt.to_homogeneous_pos(pos)
    w, v, u = (tt.to_homogeneous_vec(x) for x in (w, v, u))
    pos, w, v, u = np.broadcast_arrays(pos, w, v, u)
    if ignore_scale:
        w, v, u = (x / vc.norm(x)[:, None] for x in (w, v, u))
    return Transform(np.stack((w, v, u, pos), axis=2))


def from_perspective(*, pos=None, w=None, v=None, u=None, vol=None, ignore_scale=True):
    return to_perspective(pos=pos, w=w, v=v, u=u, vol=vol, ignore_scale=ignore_scale).inv


def random_transform(rng=None):
    rng = np.random.default_rng() if rng is None else rng
    t, pos, axis, s = rng.normal(size=(4, 3))
    angle = rng.normal()
    return rotate(pos=pos, axis=axis, angles=angle) * scale(abs(s)) * translate(t)
