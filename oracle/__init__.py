"""CPU FP64 oracle of QUMASUN's H·psi + LOBPCG path — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
package, and only as the oracle the tests compare with (or as the timed CPU baseline).  The product package
qumasun_amd never imports it.
"""
from .binding import *  # noqa: F401,F403
from .scf import *  # noqa: F401,F403,E402
from .force import *  # noqa: F401,F403,E402
from .kpoint import *  # noqa: F401,F403,E402
