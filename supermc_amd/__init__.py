"""supermc_amd -- MI355X-native implementation of superMC's per-event
initial-condition hot path (nucleon sampling, collision test, Gaussian
smearing, MC-KLN, eps_n moments) behind a C-ABI (include/smc.h).

Python here is a thin ctypes binding used by tests, bench.py and the
PyTorch calibration harness; the host side proper is the C++ facade in
supermc_amd/host/.
"""
from ._native import (Context, Profile3D, SmcError, SmcParams, SmcDerived, default_params, load,  # noqa: F401
                      EV_STRIDE, EV_EPS_RE, EV_EPS_IM, EV_EPSP_RE, EV_EPSP_IM, EV_RN, EV_XC, EV_YC,
                      EV_TOTAL, EV_DSDY, EV_NPART, EV_NCOLL, EV_B, EV_STATUS, POS_QUANTUM_FM)

__all__ = ["Context", "Profile3D", "SmcError", "SmcParams", "SmcDerived", "default_params", "load"]
