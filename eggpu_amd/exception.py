"""Exception types with the names of ``easygraph/utils/exception.py``.

When EasyGraph itself is importable its classes are reused, so callers that
catch ``easygraph.EasyGraphNotImplemented`` keep working.
"""
try:  # pragma: no cover - depends on the host environment
    from easygraph.utils.exception import (  # type: ignore
        EasyGraphError,
        EasyGraphException,
        EasyGraphNotImplemented,
        EasyGraphPointlessConcept,
    )
except Exception:  # EasyGraph not installed: standalone definitions

    class EasyGraphException(Exception):
        """Base class for exceptions in EasyGraph."""

    class EasyGraphError(EasyGraphException):
        """Exception for a serious error in EasyGraph."""

    class EasyGraphNotImplemented(EasyGraphException):
        """Raised by algorithms not implemented for a type of graph."""

    class EasyGraphPointlessConcept(EasyGraphException):
        """Raised when a concept is undefined for the graph (the null graph)."""


__all__ = ["EasyGraphException", "EasyGraphError", "EasyGraphNotImplemented", "EasyGraphPointlessConcept"]
