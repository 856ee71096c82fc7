"""Capacity convention of the reference (core/capacity.py:39-238).

Kernels never drop results silently: when an output buffer is too small they
report the required size, the host raises CapacityError or regrows to the
next power of two and calls again (core/capacity.py:175-234,
core/propagator.py:137-172).
"""


class CapacityError(ValueError):
    """Raised when an output exceeds its capacity (core/capacity.py:39-45)."""

    def __init__(self, message: str, required: int):
        super().__init__(message)
        self.required = int(required)


def next_power_of_two(n: int) -> int:
    """Growth rule of LensCapacity (core/capacity.py:175-234, base 2)."""
    n = max(int(n), 1)
    return 1 << (n - 1).bit_length()
