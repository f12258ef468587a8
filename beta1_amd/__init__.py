"""MI355X-native explicit MHD step (Rusanov + constrained transport + CFL dt).

Python view of the C-ABI in ``include/beta1_mhd.h``.  The library
``libbeta1_mhd.so`` is built in-tree by ``beta1_amd.build.build()``; importing
this package never falls back to another implementation -- if the library is
missing, :func:`lib` raises.
"""
from .capi import lib, LIB_PATH, MhdError  # noqa: F401
from .fields import FIELDS, FIELD_ID, LOCATION, NCOMP  # noqa: F401
from .solver import CaseDirDesc, HostCase, MhdContext, LunarSolver, MercuryDesc, OTDesc, WindDesc, case_parts  # noqa: F401
