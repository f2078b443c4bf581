"""FP64 CPU oracle for the neXtSIM_DG dynamics hot path — TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may
import this package, and only as the checker. The product (``nextsimdg_amd``)
never imports, links or runs anything here.
"""
from .oracle import MEB_DEFAULTS, MEB_ORDER, Oracle, OracleDyn, OracleMesh, OracleTransport, PARAM_ORDER, load, set_threads  # noqa: F401
