"""Exception hierarchy mirroring hafnium/errors.py:9-43 of the reference."""

from typing import Any


class HafniumError(Exception):
    """Base class of all exceptions raised by this package."""


class InputError(HafniumError, ValueError):
    """A user-supplied specification is invalid or inconsistent."""


class StructureError(HafniumError, ValueError):
    """Shapes or structures are inconsistent."""


class BackendError(HafniumError):
    """The device or the native library cannot perform the requested operation."""


class ConvergenceError(HafniumError):
    """An iterative procedure stopped before reaching its tolerances.

    ``last`` holds the final state (for the SCF a result with ``converged``
    false), so the caller can inspect it or restart from it.
    """

    def __init__(self, message: str, last: Any) -> None:
        super().__init__(message)
        self.last = last
