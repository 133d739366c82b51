"""The goetia names of the dBG hot path, as ``goetia.libgoetia`` exposes them
through cppyy (interface.hh / goetia.map): storages, shifters, hash models,
KmerIterator and dBG[Storage, Shifter], and around it FastxParser, the
InserterProcessor chunk driver, the filters and the compactor's segment
finder."""
from .dbg import Kmer, Shift, dBG
from .hashing import (CanLemireShifter, Canonical, DNA_SIMPLE, FwdLemireShifter, Hash,
                      KmerIteratorT as KmerIterator)
from .storage import (BitStorage, BTreeStorage, ByteStorage, PHMapStorage, SparseppSetStorage,
                      StorageTraits, count_t, get_n_primes_near_x)
from ._native import (GoetiaException, GoetiaFileException, InvalidCharacterException, InvalidSequenceException,
                      SequenceLengthException, UninitializedShifterException)
from .parsing import FastxParser, NoMoreReadsAvailable, Record, SplitPairedReader
from .processors import FilterProcessor, InserterProcessor, IntervalCounter
from .filters import DiginormFilter, StreamingSolidFilter
from .cdbg import StreamingCompactor

__all__ = ['dBG', 'Kmer', 'Shift', 'CanLemireShifter', 'FwdLemireShifter', 'Canonical', 'Hash',
           'DNA_SIMPLE', 'KmerIterator', 'BitStorage', 'ByteStorage', 'SparseppSetStorage',
           'PHMapStorage', 'BTreeStorage', 'StorageTraits', 'count_t', 'get_n_primes_near_x',
           'GoetiaException', 'InvalidSequenceException', 'SequenceLengthException',
           'UninitializedShifterException', 'GoetiaFileException', 'InvalidCharacterException',
           'FastxParser', 'NoMoreReadsAvailable', 'Record', 'SplitPairedReader', 'InserterProcessor',
           'IntervalCounter',
           'FilterProcessor',
           'StreamingSolidFilter', 'DiginormFilter', 'StreamingCompactor']
