"""smaugamd: host-side Python mirror of SMAUG's per-step MHD update path
for MI355X (workload set-ups and the ctypes binding of the HIP engine)."""
from . import configs  # noqa: F401
