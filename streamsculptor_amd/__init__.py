"""streamsculptor_amd -- MI355X-native (gfx950) stream-particle integration.

Drop-in for the stream-particle path of the reference ``streamsculptor`` package: ``Potential``
subclasses with potential / gradient / acceleration, ``integrate_orbit``, ``gen_stream_ics`` /
``gen_stream_vmapped`` and the ``fields`` auxiliary-ODE interface, computed by hand-written HIP
kernels in ``libssc_mi355x.so`` (C-ABI: include/ssc_mi355x.h).
"""
from ._lib import G_USYS, ExtensionMissing  # noqa: F401
from .main import Potential, Solution, SourceSet, UnitSystem, usys, dimensionless  # noqa: F401
from . import solvers  # noqa: F401
from .solvers import Leapfrog, Dopri5, Dopri8  # noqa: F401
from . import potential  # noqa: F401
from .streamhelpers import (gen_stream_ics_pert, gen_stream_vmapped_with_pert,  # noqa: F401
                            gen_stream_vmapped_with_pert_fixed_prog)

__version__ = "0.1.0"
