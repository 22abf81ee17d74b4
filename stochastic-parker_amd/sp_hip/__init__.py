"""MI355X Parker pseudo-particle pusher: Python host bindings of libsp_hip.so."""
from ._abi import SP_PREC_F32, SP_PREC_F64, SpError, load  # noqa: F401
from .pusher import ParkerPusher, empty_particles, make_config  # noqa: F401
