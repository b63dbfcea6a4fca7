"""Compute-class plumbing mirroring freud/util.py:11-168."""

from functools import wraps

import numpy as np

import freud_amd.box

NO_DEFAULT_QUERY_ARGS_MESSAGE = (
    "The {} class does not provide default query arguments. Query arguments "
    "or a neighbor list must be provided to this compute method."
)


class _Compute:
    """Guards computed properties until ``compute`` has run (freud/util.py:11-83)."""

    def __init__(self):
        self._called_compute = False

    def __getattribute__(self, attr):
        attribute = object.__getattribute__(self, attr)
        if attr == "compute":
            compute = attribute

            @wraps(compute)
            def compute_wrapper(*args, **kwargs):
                return_value = compute(*args, **kwargs)
                self._called_compute = True
                return return_value

            return compute_wrapper
        return attribute

    @staticmethod
    def _computed_property(prop):
        @property
        @wraps(prop)
        def wrapper(self, *args, **kwargs):
            if not getattr(self, "_called_compute", False):
                msg = "Property not computed. Call compute first."
                raise AttributeError(msg)
            return prop(self, *args, **kwargs)

        return wrapper

    def __str__(self):
        return repr(self)


def _convert_array(array, shape=None, dtype=np.float32, requirements=("C",), allow_copy=True):
    """freud/util.py:86-141."""
    if array is None:
        return np.empty(shape, dtype=dtype)
    array = np.asarray(array)
    return_arr = np.require(array, dtype=dtype, requirements=requirements)
    if not allow_copy and return_arr is not array:
        msg = (
            "The provided output array must have dtype "
            f"{dtype}, and have the following array flags: {', '.join(requirements)}."
        )
        raise ValueError(msg)
    if shape is not None:
        if return_arr.ndim != len(shape):
            msg = f"array.ndim = {return_arr.ndim}; expected ndim = {len(shape)}"
            raise ValueError(msg)
        for i, s in enumerate(shape):
            if s is not None and return_arr.shape[i] != s:
                shape_str = "(" + ", ".join(str(i) if i is not None else "..." for i in shape) + ")"
                msg = f"array.shape= {return_arr.shape}; expected shape = {shape_str}"
                raise ValueError(msg)
    return return_arr


def _convert_box(box, dimensions=None):
    """freud/util.py:144-168."""
    if not isinstance(box, freud_amd.box.Box):
        box = freud_amd.box.Box.from_box(box)
    if dimensions is not None and box.dimensions != dimensions:
        msg = f"The box must be {dimensions}-dimensional."
        raise ValueError(msg)
    return box
