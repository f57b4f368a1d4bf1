from . import gauge  # noqa: F401
from . import fermion  # noqa: F401
