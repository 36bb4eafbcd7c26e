"""Small helpers shared by the geometry classes."""
from numbers import Integral


def _up_slice(key):
    if isinstance(key, Integral):
        return slice(key, None) if key == -1 else slice(key, key + 1)
    return key


def slice_interval(left, right, length, key):
    """Slice the interval ``[left, right]`` split into ``length`` pixels.

    Returns ``(new_left, new_right, new_length, new_pixel_size)``. A step ``s``
    multiplies the pixel size by ``s`` and keeps each selected pixel centred on
    the original pixel centre, as tomosipo ``utils.py:45-119`` does, so binned
    pixels may extend beyond the original interval.
    """
    pixel_size = 1 if length == 0 else (right - left) / length
    start, stop, step = _up_slice(key).indices(length)
    new_len = max(0, (stop - start + step - 1) // step)
    stop = max(start, start + (new_len - 1) * step + 1)
    L = left + start * pixel_size + 0.5 * pixel_size * (1 - step)
    R = left + stop * pixel_size + 0.5 * pixel_size * (step - 1)
    return (L, R, new_len, pixel_size * step)
