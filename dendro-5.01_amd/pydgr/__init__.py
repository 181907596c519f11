"""Python side of the MI355X BSSN RHS library: ctypes binding of the C ABI
(include/dgr_bssn.h), block tables and seeded synthetic data."""
from .native import (ETS, ETS_RK3, ETS_RK4, NUM_VARS, VAR_NAMES, BSSNParams, Block, DgrError, Plan, RK4,  # noqa: F401
                     Schedule, TwoPunctures,
                     default_schedule, lib, make_params, rhs_blocks, tp_params)
from . import layout, data  # noqa: F401
