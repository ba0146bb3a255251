"""MI355X-native T-FlowS Krylov linear-solver path (CG / BiCG / CGS + Prec1).

The directory name carries hyphens, so it is loaded as the package
``tflows_amd`` by :func:`load` in ``__graft_entry__`` and ``tests/conftest``.
"""
from . import meshgen  # noqa: F401
from . import krylov  # noqa: F401,E402
from . import adjoint_harness  # noqa: F401,E402
