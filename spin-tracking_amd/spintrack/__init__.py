"""spintrack -- MI355X (gfx950) spin-tracking hot path, host-side interface.

The product is the C ABI in include/spintrack.h, implemented by
``libspintrack.so`` (HIP kernels for gfx950).  This package mirrors the
reference's host interface on top of it:

* ``options``    -- the 440-byte options block, params.txt grammar, tc/sqrtKT_m
* ``particle``   -- the ``particle`` class (init / runSimulation / outputData)
* ``logger``     -- the binary snapshot format (writer and reader)
* ``simulation`` -- the ``mainAnalysis`` run loop
* ``floquet``    -- host-side Floquet routines of the spectral path (type 6)
"""
from ._lib import (VARIANT_EXACT, VARIANT_FAST, Coords, Ext, Options, SpinTrackError,  # noqa: F401
                   load)
from .options import (HELIUM3, NEUTRON, default_options, derive, n_intervals,  # noqa: F401
                      output_times, parse_params, read_params, set_fields)
from .particle import COORDS, Particle, snapshot_block_dtype, xyz  # noqa: F401
from .logger import BinaryLogger, read_output  # noqa: F401
from . import adjoint, configs, floquet, rng, simulation  # noqa: F401
