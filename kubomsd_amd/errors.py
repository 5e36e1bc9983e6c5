"""Error hierarchy (reference src/kubo_msd/errors.py:1-8): every failure is a
ValueError subclass so callers written against the reference catch it."""


class KuboMSDError(ValueError):
    """Invalid input or failed numerical acceptance; never silently repaired."""


class BackendUnavailable(KuboMSDError):
    pass


class NumericalError(KuboMSDError):
    pass


class ResourceBlocked(NumericalError):
    """Device memory budget refused an allocation (message starts RESOURCE_BLOCKED)."""
