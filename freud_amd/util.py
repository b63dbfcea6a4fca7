"""Host-side helpers shared by the compute classes.

Behaviour follows freud's compute classes (freud/util.py:11-168 of freud
3.5.0): result properties raise ``AttributeError`` until ``compute`` has
returned once; array arguments are coerced to C-contiguous arrays of a given
dtype, copying only when needed, and their shapes are checked; box-like
arguments become :class:`freud_amd.box.Box`.
"""

import functools

import numpy as np

import freud_amd.box

# Text of the error raised by compute classes that need explicit neighbours
# (freud/errors.py:15-18).
NO_DEFAULT_QUERY_ARGS_MESSAGE = (
    "The {} class does not provide default query arguments. Query arguments "
    "or a neighbor list must be provided to this compute method."
)


class computed_property:
    """A read-only property whose getter runs only after ``compute``.

    The owning object records a completed ``compute`` call in
    ``_called_compute`` (see :class:`_Compute`); reading the property before
    that raises ``AttributeError``, as freud's properties do.
    """

    def __init__(self, getter):
        self._getter = getter
        functools.update_wrapper(self, getter)

    def __set_name__(self, owner, name):
        self._name = name

    def __get__(self, obj, objtype=None):
        if obj is None:
            return self
        if not obj.__dict__.get("_called_compute", False):
            raise AttributeError(f"Property {getattr(self, '_name', '')!s} not computed. Call compute first.")
        return self._getter(obj)

    def __set__(self, obj, value):
        raise AttributeError("computed properties are read-only")


class _Compute:
    """Base of the compute classes.

    Every ``compute`` method defined by a subclass is wrapped when the class is
    created: after it returns without raising, the instance is marked as
    computed, which unlocks its :class:`computed_property` attributes.
    """

    def __init__(self):
        self._called_compute = False

    def __init_subclass__(cls, **kwargs):
        super().__init_subclass__(**kwargs)
        fn = cls.__dict__.get("compute")
        if fn is None or getattr(fn, "_marks_computed", False):
            return

        @functools.wraps(fn)
        def compute(self, *args, **kw):
            out = fn(self, *args, **kw)
            self.__dict__["_called_compute"] = True
            return out

        compute._marks_computed = True
        cls.compute = compute

    def __str__(self):
        return repr(self)


def _check_shape(arr, shape):
    # the messages of freud/util.py:126-139, which callers match on
    if len(shape) != arr.ndim:
        raise ValueError(f"array.ndim = {arr.ndim}; expected ndim = {len(shape)}")
    for want, have in zip(shape, arr.shape):
        if want is not None and want != have:
            expect = "(" + ", ".join("..." if s is None else str(s) for s in shape) + ")"
            raise ValueError(f"array.shape= {arr.shape}; expected shape = {expect}")


def _convert_array(array, shape=None, dtype=np.float32, requirements=("C",), allow_copy=True):
    """``array`` as an ndarray of ``dtype`` with the given flags (a copy only
    when the input does not already qualify); ``shape`` entries of None match
    any length.  None gives an uninitialised array of ``shape``.  With
    ``allow_copy=False`` an input that would need a copy is rejected."""
    if array is None:
        return np.empty(shape, dtype=dtype)
    src = np.asarray(array)
    out = np.require(src, dtype=dtype, requirements=requirements)
    if out is not src and not allow_copy:
        raise ValueError(f"the output array must already be {np.dtype(dtype).name} with flags {', '.join(requirements)}")
    if shape is not None:
        _check_shape(out, shape)
    return out


def _convert_box(box, dimensions=None):
    """A :class:`freud_amd.box.Box` from any box-like value; with
    ``dimensions`` (2 or 3) the box must have that many."""
    out = box if isinstance(box, freud_amd.box.Box) else freud_amd.box.Box.from_box(box)
    if dimensions is not None and out.dimensions != dimensions:
        raise ValueError(f"The box must be {dimensions}-dimensional.")
    return out
