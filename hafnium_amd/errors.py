"""Exception classes of the package.

The class names, their bases and the ``last`` attribute of
``ConvergenceError`` are the reference's exception contract
(hafnium/errors.py:9-43), kept so that callers written against the
reference catch the same types.
"""

from typing import Any


class HafniumError(Exception):
    """Root of every exception this package raises."""


class InputError(HafniumError, ValueError):
    """Arguments that are malformed, out of range or contradict each other (shapes, counts, options)."""


class StructureError(HafniumError, ValueError):
    """Array shapes or block layouts that do not fit together."""


class BackendError(HafniumError):
    """The device or the native library cannot perform the requested operation."""


class ConvergenceError(HafniumError):
    """An iteration hit its limit with its tolerances unmet.

    The attribute ``last`` carries the state reached (for the SCF, a result
    whose ``converged`` is false): it can be examined, or passed back as the
    start of another run.
    """

    def __init__(self, message: str, last: Any) -> None:
        super().__init__(message)
        self.last = last
