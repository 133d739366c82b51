"""ctypes binding of libgoetia_amd.so (the C ABI in include/goetia_amd.h).

The library is built in-tree by ``python -m goetia_amd.build``.  There is no
fallback: if the library is missing, or a device entry point runs without a
GPU, the call raises.
"""
import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libgoetia_amd.so')
HEADER = os.path.join(os.path.dirname(HERE), 'include', 'goetia_amd.h')

GT_OK = 0
GT_ERR_INVALID_ARG = 1
GT_ERR_SEQUENCE_LENGTH = 2
GT_ERR_OUT_OF_MEMORY = 3
GT_ERR_DEVICE = 4
GT_ERR_UNSUPPORTED = 5
GT_ERR_NO_DEVICE = 6
GT_ERR_IO = 7
GT_ERR_INTERNAL = 8
GT_ERR_INVALID_CHARACTER = 9
GT_ERR_PAIRING = 10

BIT, BYTE, EXACT = 1, 2, 3
MODE_RAW, MODE_SKIP_N = 0, 1


class GoetiaException(Exception):
    """goetia::GoetiaException (goetia.hh:140-153)."""


class SequenceLengthException(GoetiaException):
    """goetia::SequenceLengthException (exceptions.hh:23-28)."""


class InvalidSequenceException(GoetiaException):
    """goetia::InvalidSequenceException (raised by hash_base on short input)."""


class UninitializedShifterException(GoetiaException):
    """goetia::UninitializedShifterException (hashshifter.hh:133-145)."""


class InvalidCharacterException(GoetiaException):
    """goetia::InvalidCharacterException (alphabets.hh:41-55)."""


class GoetiaFileException(GoetiaException, OSError):
    """goetia::GoetiaFileException (goetia.hh): a file could not be read."""


class DeviceError(GoetiaException):
    """The GPU path failed or no GPU is available."""


class UnsupportedError(GoetiaException):
    pass


_ERRORS = {
    GT_ERR_INVALID_ARG: ValueError,
    GT_ERR_SEQUENCE_LENGTH: SequenceLengthException,
    GT_ERR_OUT_OF_MEMORY: MemoryError,
    GT_ERR_DEVICE: DeviceError,
    GT_ERR_UNSUPPORTED: UnsupportedError,
    GT_ERR_NO_DEVICE: DeviceError,
    GT_ERR_IO: GoetiaFileException,
    GT_ERR_INTERNAL: GoetiaException,
    GT_ERR_INVALID_CHARACTER: InvalidCharacterException,
    GT_ERR_PAIRING: GoetiaException,
}

u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
i64p = C.POINTER(C.c_int64)
i16p = C.POINTER(C.c_int16)
vp = C.c_void_p
st_p = C.c_void_p
u16 = C.c_uint16
u32 = C.c_uint32
u64 = C.c_uint64
cint = C.c_int

# name -> argument types (all return int status, except gt_last_error)
SIGNATURES = {
    'gt_abi_version': [],
    'gt_device_init': [cint],
    'gt_device_synchronize': [],
    'gt_primes_near_x': [u32, u64, u64p, u32p],
    'gt_storage_new': [cint, u64p, u32, u64, C.POINTER(vp)],
    'gt_storage_free': [st_p],
    'gt_storage_clone': [st_p, C.POINTER(vp)],
    'gt_storage_reset': [st_p],
    'gt_storage_kind': [st_p, C.POINTER(cint)],
    'gt_storage_n_tables': [st_p, u32p],
    'gt_storage_tablesizes': [st_p, u64p],
    'gt_storage_table_bytes': [st_p, u32, u64p],
    'gt_storage_counters': [st_p, u64p, u64p],
    'gt_storage_set_counters': [st_p, u64, u64],
    'gt_storage_set_use_bigcount': [st_p, C.c_int],
    'gt_storage_get_use_bigcount': [st_p, C.POINTER(C.c_int)],
    'gt_storage_bigcounts': [st_p, vp, vp, u64, u64p],
    'gt_storage_set_bigcounts': [st_p, vp, vp, u64],
    'gt_storage_read_table': [st_p, u32, vp],
    'gt_storage_write_table': [st_p, u32, vp],
    'gt_storage_update_from': [st_p, st_p],
    'gt_storage_exact_keys': [st_p, vp, u64, u64p],
    'gt_storage_save_async': [st_p, vp, vp, vp],
    'gt_storage_load_async': [st_p, vp, vp, vp],
    'gt_storage_save': [st_p, C.c_char_p, u16],
    'gt_storage_load': [st_p, C.c_char_p, C.POINTER(u16)],
    'gt_oxli_write': [C.c_char_p, cint, u16, u32, u64p, u64, vp],
    'gt_oxli_read_header': [C.c_char_p, cint, C.POINTER(u16), u32p, u64p, u64p],
    'gt_oxli_read': [C.c_char_p, cint, vp],
    'gt_kmer_offsets_dev': [u16, vp, u64, vp, vp],
    'gt_hash_reads_dev': [u16, cint, vp, vp, u64, u64, vp, vp, vp, vp],
    'gt_hash_reads': [u16, cint, vp, vp, u64, vp, vp, vp],
    'gt_shifter_hash_base': [u16, vp, u64p, u64p],
    'gt_shifter_shift_right': [u16, u64p, u64p, C.c_uint8, C.c_uint8],
    'gt_shifter_shift_left': [u16, u64p, u64p, C.c_uint8, C.c_uint8],
    'gt_dbg_insert_dev': [st_p, u16, cint, cint, vp, vp, u64, u64, vp, vp, vp],
    'gt_dbg_insert': [st_p, u16, cint, cint, vp, vp, u64, vp, vp, u64p],
    'gt_dbg_insert_noalloc_dev': [st_p, u16, cint, cint, vp, vp, u64, u64, vp, vp, vp],
    'gt_dbg_query_dev': [st_p, u16, cint, cint, vp, vp, u64, u64, vp, vp, vp],
    'gt_dbg_query': [st_p, u16, cint, cint, vp, vp, u64, vp, vp],
    'gt_dbg_insert_and_query_dev': [st_p, u16, cint, cint, vp, vp, u64, u64, vp, vp],
    'gt_dbg_insert_and_query': [st_p, u16, cint, cint, vp, vp, u64, vp],
    'gt_dbg_degrees_dev': [st_p, u16, cint, cint, cint, vp, vp, u64, u64, vp, vp, vp, vp],
    'gt_dbg_degrees': [st_p, u16, cint, cint, cint, vp, vp, u64, vp, vp, vp],
    'gt_dbg_neighbors_dev': [st_p, u16, cint, cint, cint, vp, vp, u64, u64, vp, vp, vp, vp, vp],
    'gt_dbg_neighbors': [st_p, u16, cint, cint, cint, vp, vp, u64, vp, vp, vp, vp],
    'gt_find_new_segments': [st_p, u16, cint, vp, vp, u64, vp, u64, cint, vp, u64, vp, vp, vp, vp, u64, vp, vp, u64,
                             vp, vp],
    'gt_segments_from_view': [u16, vp, u64, vp, vp, vp, vp, vp, u64, cint, vp, u64, vp, vp, u64, vp, vp, u64, vp, vp],
    'gt_insert_hashes_dev': [st_p, vp, u64, vp, vp],
    'gt_insert_hashes': [st_p, vp, u64, vp],
    'gt_query_hashes_dev': [st_p, vp, u64, vp, vp],
    'gt_query_hashes': [st_p, vp, u64, vp],
    'gt_insert_and_query_hashes': [st_p, vp, u64, vp],
    'gt_tune_set': [cint, cint, cint],
    'gt_tune_get': [cint, C.POINTER(cint), C.POINTER(cint)],
    'gt_autotune': [st_p, u16, cint, cint, vp, vp, u64, u64, C.POINTER(cint), C.POINTER(cint), vp],
    'gt_pipeline_new': [st_p, u16, cint, cint, u64, u64, C.POINTER(vp)],
    'gt_pipeline_submit': [vp, vp, vp, u64, vp],
    'gt_pipeline_flush': [vp],
    'gt_pipeline_free': [vp],
    'gt_workspace_new': [C.POINTER(vp)],
    'gt_workspace_free': [vp],
    'gt_dbg_insert_ws_dev': [st_p, vp, u16, cint, cint, vp, vp, u64, u64, vp, vp, vp],
    'gt_dbg_query_ws_dev': [st_p, vp, u16, cint, cint, vp, vp, u64, u64, vp, vp, vp],
    'gt_server_new': [cint, u64p, u32, u16, cint, cint, u32, u64, u32, cint, C.POINTER(vp)],
    'gt_server_submit': [vp, u32, vp, vp, u64, vp, u64p],
    'gt_server_wait': [vp, u64, C.POINTER(C.c_double)],
    'gt_server_drain': [vp],
    'gt_server_get_stats': [vp, vp],
    'gt_server_stream_table': [vp, u32, u32, vp, vp],
    'gt_server_free': [vp],
    'gt_profile_enable': [cint],
    'gt_profile_reset': [],
    'gt_profile_get': [C.c_char_p, u64p, C.POINTER(C.c_double)],
    'gt_gen_genome': [u64, u64, vp],
    'gt_gen_reads': [u64, vp, u64, u64, vp, u32, u32, C.c_double, C.c_double, vp, vp, u64p],
    'gt_solid_filter_dev': [st_p, u16, cint, cint, vp, vp, u64, u64, C.c_float, u32, vp, vp],
    'gt_solid_filter': [st_p, u16, cint, cint, vp, vp, u64, C.c_float, u32, vp],
    'gt_diginorm_dev': [st_p, u16, cint, cint, vp, vp, u64, u64, u32, vp, vp],
    'gt_diginorm': [st_p, u16, cint, cint, vp, vp, u64, u32, vp],
    'gt_membench_random': [vp, u64, u64, cint, u64, vp, u64p],
    'gt_fastx_open': [C.c_char_p, cint, u32, C.POINTER(vp)],
    'gt_fastx_next_batch': [vp, u64, u64, vp, vp, vp, u64, vp, u64p],
    'gt_fastx_next_batch_qual': [vp, u64, u64, vp, vp, vp, vp, vp, u64, vp, u64p],
    'gt_fastx_stats': [vp, u64p, u64p, u64p, C.POINTER(cint)],
    'gt_fastx_close': [vp],
    'gt_inserter_new': [st_p, C.c_uint16, cint, cint, u64, u64, C.POINTER(vp)],
    'gt_inserter_advance': [vp, vp, u64p, u64p, C.POINTER(cint)],
    'gt_inserter_stats': [vp, u64p, u64p, u64p],
    'gt_inserter_free': [vp],
    'gt_pairs_open': [C.c_char_p, C.c_char_p, cint, u32, cint, C.POINTER(vp)],
    'gt_pairs_next': [vp, vp, u64, vp, vp, u64, vp, C.POINTER(cint)],
    'gt_pairs_stats': [vp, u64p, u64p, C.POINTER(cint)],
    'gt_pairs_close': [vp],
    'gt_inserter_advance_pairs': [vp, vp, u64p, u64p, C.POINTER(cint)],
}

_lib = None


def header_symbols():
    """Names of the entry points declared in include/goetia_amd.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r'^(?:int|const char\*)\s+(gt_\w+)\(', text, re.M)))


def load():
    """Loads libgoetia_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f'{LIB_PATH} is missing: build it with `python -m goetia_amd.build` '
                          '(goetia_amd has no CPU fallback)')
    # torch-ROCm ships its own HIP runtime (torch/lib/libamdhip64.so, soname
    # libamdhip64.so.7) and links it by the unversioned name.  Loaded first,
    # it satisfies this library's libamdhip64.so.7 dependency, so the process
    # holds one HIP runtime.  Loaded the other way round, the dynamic linker
    # maps /opt/rocm's runtime for us and then a second copy for torch: two
    # HSA runtimes in one process, and torch's later device initialisation
    # fails ("No HIP GPUs are available").
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    for name in ('gt_last_error', 'gt_server_last_error'):
        getattr(lib, name).restype = C.c_char_p
        getattr(lib, name).argtypes = []
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = C.c_int
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc, what=''):
    if rc != GT_OK:
        err = load().gt_server_last_error if what.startswith('gt_server') else load().gt_last_error
        msg = err().decode(errors='replace')
        exc = _ERRORS.get(rc, GoetiaException)
        raise exc(f'{what}: {msg}' if what else msg)


def call(name, *args):
    rc = getattr(load(), name)(*args)
    check(rc, name)
    return rc


def ptr(arr):
    """Address of a numpy array or torch tensor (None -> NULL)."""
    if arr is None:
        return None
    if hasattr(arr, 'data_ptr'):
        return arr.data_ptr()
    return arr.ctypes.data
