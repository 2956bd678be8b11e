"""Exception types of the backend boundary.

``CapabilityError`` mirrors ``qarp.errors.CapabilityError`` (a ``ValueError``,
src/qarp/errors.py), the translation target of qarpx::capability_error
(src/qarpx/python/bindings.cpp:272-290)."""


class CapabilityError(ValueError):
    """A request the backend cannot honour (e.g. a symbolic gate with no
    adjoint generator rule)."""


class BackendUnavailable(RuntimeError):
    """The HIP library is missing or no GPU is visible."""
