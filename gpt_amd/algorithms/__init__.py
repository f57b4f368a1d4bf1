from . import inverter  # noqa: F401
