"""Row-wise ``(n, 3)`` vector algebra and 4x4 homogeneous transforms.

Same operations as tomosipo's ``vector_calc.py`` (``to_vec`` ``:19``,
``cross_product`` ``:117``, ``intersect`` ``:128``, ``dot``/``norm``
``:142-148``, ``matrix_transform`` ``:195``, ``invert_transformation_matrix``
``:211``); rows broadcast when one side has a single row.
"""
import numpy as np

from .types import to_vec, to_homogeneous, to_homogeneous_vec

to_homogeneous_point = lambda x: to_homogeneous(x, 1.0)  # noqa: E731


def broadcast_lengths(len_a, len_b):
    if len_a == 1 or len_b == 1 or len_a == len_b:
        return max(len_a, len_b)
    raise ValueError(f"Lengths {len_a} and {len_b} cannot be broadcast.")


def check_same_shapes(*args):
    shapes = [a.shape for a in args]
    if min(shapes) != max(shapes):
        raise ValueError(f"Not all arguments are the same shape. Got: {shapes}")


def dot(x, y):
    return np.sum(np.asarray(x) * np.asarray(y), axis=-1)


def squared_norm(x):
    return dot(x, x)


def norm(x):
    return np.sqrt(squared_norm(x))


def cross_product(x, y):
    x, y = np.broadcast_arrays(to_vec(x), to_vec(y))
    return np.cross(x, y)


def intersect(v_origin, v_direction, plane_origin, plane_normal):
    """Intersection of the lines ``o + t d`` with the planes through ``p`` with normal ``n``."""
    v_origin, v_direction, plane_origin, plane_normal = np.broadcast_arrays(
        to_vec(v_origin), to_vec(v_direction), to_vec(plane_origin), to_vec(plane_normal)
    )
    with np.errstate(divide="ignore", invalid="ignore"):
        t = dot(plane_origin - v_origin, plane_normal) / dot(v_direction, plane_normal)
    return v_origin + t[:, None] * v_direction


def to_matrices(M):
    M = np.array(M, dtype=np.float64, copy=True)
    if M.ndim == 2:
        M = M[None]
    if M.ndim != 3 or M.shape[1:] != (4, 4):
        raise ValueError(f"Expected (n, 4, 4) transformation matrices. Got {M.shape}.")
    return M


def matrix_transform(M, x):
    """Apply ``M`` (n,4,4) to homogeneous rows ``x`` (n,4), broadcasting on n."""
    M = to_matrices(M)
    n = broadcast_lengths(len(M), len(x))
    M = np.broadcast_to(M, (n, 4, 4))
    x = np.broadcast_to(x, (n, 4))
    return np.einsum("nij,nj->ni", M, x)


def matrix_matrix_transform(M1, M2):
    M1, M2 = to_matrices(M1), to_matrices(M2)
    n = broadcast_lengths(len(M1), len(M2))
    return np.matmul(np.broadcast_to(M1, (n, 4, 4)), np.broadcast_to(M2, (n, 4, 4)))


def invert_transformation_matrix(M):
    return np.linalg.inv(to_matrices(M))


def from_homogeneous(x):
    return np.array(x[:, :3], copy=True)
