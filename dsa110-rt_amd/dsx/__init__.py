"""dsx -- MI355X (gfx950) search-node fast path for DSA-110.

4-bit voltage unpack -> int8-MFMA correlator (exact int32 visibilities) ->
nearest-cell gridder + C2R FFT imager -> subband dedispersion fused with the
boxcar bank and float64 running S/N statistics -> candidate decode.

Stage APIs keep the names, arguments and layouts of dsa110-rt's Python
stages; the work runs in hand-written HIP kernels of ``libdsx.so`` through
the C ABI in ``include/dsx.h``.  There is no CPU fallback.
"""
from .config import NodeConfig, small_node_config, search_node_config, search_node_short_config
from .detector import Candidate, CandidateFlags, boxcar_via_cumsum, multi_boxcar_argmax

__all__ = [
    "NodeConfig", "small_node_config", "search_node_config", "search_node_short_config",
    "Candidate", "CandidateFlags", "boxcar_via_cumsum", "multi_boxcar_argmax",
]

__version__ = "0.1.0"
