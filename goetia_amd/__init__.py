"""goetia_amd: MI355X-native streaming de Bruijn graph hot path.

Canonical rolling k-mer hashing feeding Bloom (BitStorage), count-min
(ByteStorage) and exact k-mer set storages, with goetia's dBG API, running
as hand-written HIP kernels for gfx950 behind a C ABI (include/goetia_amd.h).
"""
from . import libgoetia
from ._native import DeviceError, GoetiaException, MODE_RAW, MODE_SKIP_N
from .dbg import dBG
from .hashing import CanLemireShifter, FwdLemireShifter
from .reads import DeviceReads, ReadBatch
from .storage import BitStorage, ByteStorage, SparseppSetStorage

__version__ = '0.1.0'
