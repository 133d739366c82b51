"""OXLI v4 files against goetia's own (tests/golden/data/oxli/, written by
goetia's compiled storages through tests/golden/make_oxli_golden.py, with
what goetia's loader read back in tests/golden/oxli_ref.json).

Host (CPU): the library's reader gives goetia's header fields and tables for
every file, plain and gzip; its writer, given those tables, writes the same
bytes, compressed files included (goetia picks gzip for count-min files
named *.gz; Bloom files are never compressed); the host reader skips a
count-min file's bigcount pairs;
where the compiled reference is built, goetia's loader reads files the
library wrote.
GPU: a device storage filled as goetia's was saves the same bytes, and loads
goetia's files (taking their table sizes) with the same tables and counters.
"""
import ctypes as C
import hashlib
import json
import os
import shutil

import numpy as np
import pytest

from conftest import GOLDEN_DIR, read_fasta, read_fastq

OXLI_DIR = os.path.join(GOLDEN_DIR, 'data', 'oxli')
META = json.load(open(os.path.join(GOLDEN_DIR, 'oxli_ref.json')))
FILES = sorted(META)
REF_SO = os.path.join(os.path.dirname(os.path.dirname(GOLDEN_DIR)), 'oracle', '_ref', 'libgoetia_ref.so')


def sha(b):
    return hashlib.sha256(b).hexdigest()


@pytest.mark.parametrize('name', FILES)
def test_read_goetia_file(name):
    from goetia_amd import oxli
    m = META[name]
    k, sizes, occ, tables = oxli.read_file(os.path.join(OXLI_DIR, name), m['kind'])
    assert (k, sizes, occ) == (m['ksize'], m['tablesizes'], m['occupied'])
    assert [sha(t.tobytes()) for t in tables] == m['table_sha256']


@pytest.mark.parametrize('name', FILES)
def test_write_matches_goetia_bytes(name, tmp_path):
    from goetia_amd import oxli
    m = META[name]
    k, sizes, occ, tables = oxli.read_file(os.path.join(OXLI_DIR, name), m['kind'])
    out = tmp_path / name
    oxli.write_file(out, m['kind'], k, sizes, occ, tables)
    data = out.read_bytes()
    assert (data[:2] == b'\x1f\x8b') == m['gzip']
    assert sha(data) == m['file_sha256']


def _patched(tmp_path, name, offset, value):
    raw = bytearray(open(os.path.join(OXLI_DIR, name), 'rb').read())
    raw[offset:offset + len(value)] = value
    p = tmp_path / ('x_' + name)
    p.write_bytes(bytes(raw))
    return p


def test_bigcount_pairs_skipped_by_host_reader(tmp_path):
    """A count-min file with use_bigcount set and a bigcount pair reads with
    the same tables; a pair count with no pair after it is a file error."""
    from goetia_amd import _native as N
    from goetia_amd import oxli
    m = META['counts_k21.oxli']
    base = oxli.read_file(os.path.join(OXLI_DIR, 'counts_k21.oxli'), N.BYTE)
    p = _patched(tmp_path, 'counts_k21.oxli', 6, b'\x01')       # use_bigcount = 1
    got = oxli.read_file(p, N.BYTE)
    assert got[:3] == base[:3] and all(np.array_equal(a, b) for a, b in zip(got[3], base[3]))
    size = os.path.getsize(os.path.join(OXLI_DIR, 'counts_k21.oxli'))
    p = _patched(tmp_path, 'counts_k21.oxli', size - 8, (1).to_bytes(8, 'little'))   # one bigcount pair
    with pytest.raises(N.GoetiaFileException):
        oxli.read_file(p, N.BYTE)
    with open(p, 'ab') as f:
        f.write((12345).to_bytes(8, 'little') + (300).to_bytes(2, 'little'))
    got = oxli.read_file(p, N.BYTE)
    assert all(np.array_equal(a, b) for a, b in zip(got[3], base[3]))
    assert m['kind'] == N.BYTE


def test_malformed_files(tmp_path):
    from goetia_amd import _native as N
    from goetia_amd import oxli
    with pytest.raises(N.GoetiaFileException):
        oxli.read_file(tmp_path / 'missing.oxli', N.BIT)
    with pytest.raises(N.GoetiaFileException):
        oxli.read_file(_patched(tmp_path, 'bloom_k21.oxli', 0, b'OXLJ'), N.BIT)
    with pytest.raises(N.GoetiaFileException):
        oxli.read_file(_patched(tmp_path, 'bloom_k21.oxli', 4, b'\x03'), N.BIT)      # version 3
    with pytest.raises(N.GoetiaFileException):
        oxli.read_file(os.path.join(OXLI_DIR, 'bloom_k21.oxli'), N.BYTE)             # Bloom file as counts
    raw = open(os.path.join(OXLI_DIR, 'counts_k21.oxli'), 'rb').read()
    (tmp_path / 'cut.oxli').write_bytes(raw[:len(raw) // 2])
    with pytest.raises(N.GoetiaFileException):
        oxli.read_file(tmp_path / 'cut.oxli', N.BYTE)
    shutil.copy(os.path.join(OXLI_DIR, 'counts_k21.oxli'), tmp_path / 'plain_named.gz')
    k, sizes, _, _ = oxli.read_file(tmp_path / 'plain_named.gz', N.BYTE)   # zlib reads plain data too
    assert k == 21


@pytest.mark.skipif(not os.path.exists(REF_SO), reason='compiled reference (oracle/_ref) not built here')
@pytest.mark.parametrize('name', FILES)
def test_goetia_reads_library_files(name, tmp_path):
    from goetia_amd import oxli
    m = META[name]
    k, sizes, occ, tables = oxli.read_file(os.path.join(OXLI_DIR, name), m['kind'])
    # change the tables so the file is not a copy of goetia's
    tables = [np.roll(t, 7) for t in tables]
    out = tmp_path / name
    oxli.write_file(out, m['kind'], k + 2, sizes, occ + 1, tables)
    L = C.CDLL(REF_SO)
    L.ref_load_file.restype = C.c_uint64
    L.ref_load_file.argtypes = [C.c_int, C.c_char_p, C.POINTER(C.c_uint16), C.POINTER(C.c_uint32), C.c_void_p,
                                C.POINTER(C.c_uint64), C.c_void_p, C.c_uint64]
    kk, n, o = C.c_uint16(), C.c_uint32(), C.c_uint64()
    sz = np.zeros(16, np.uint64)
    buf = np.zeros(1 << 20, np.uint8)
    w = L.ref_load_file(m['kind'], str(out).encode(), C.byref(kk), C.byref(n), sz.ctypes.data, C.byref(o),
                        buf.ctypes.data, len(buf))
    assert w == sum(len(t) for t in tables)
    assert (kk.value, [int(x) for x in sz[:n.value]], o.value) == (k + 2, sizes, occ + 1)
    np.testing.assert_array_equal(buf[:w], np.concatenate(tables))


def _fill_like_goetia(m):
    from goetia_amd import libgoetia as lg
    path = os.path.join(GOLDEN_DIR, 'data', m['input'])
    reads = read_fasta(path) if m['input'].endswith('.fa') else read_fastq(path)
    cls = lg.BitStorage if m['kind'] == 1 else lg.ByteStorage
    sh = lg.CanLemireShifter if m['canonical'] else lg.FwdLemireShifter
    store = cls.build(m['max_table'], m['n_tables'])
    g = lg.dBG[cls, sh].build(store, m['K'])
    for _ in range(m['repeats']):
        g.insert_sequences(reads)
    return g, store


@pytest.mark.gpu
@pytest.mark.parametrize('name', FILES)
def test_device_save_matches_goetia(name, tmp_path):
    g, store = _fill_like_goetia(META[name])
    out = tmp_path / name
    g.save(str(out))
    assert sha(out.read_bytes()) == META[name]['file_sha256']


@pytest.mark.gpu
@pytest.mark.parametrize('name', FILES)
def test_device_load_goetia_file(name):
    from goetia_amd import libgoetia as lg
    m = META[name]
    cls = lg.BitStorage if m['kind'] == 1 else lg.ByteStorage
    store = cls.build(1000, 1)                # other sizes: the load takes the file's
    k = store.load(os.path.join(OXLI_DIR, name))
    assert k == m['ksize']
    assert store.get_tablesizes() == m['tablesizes']
    assert store.n_occupied() == m['occupied']
    assert [sha(t.tobytes()) for t in store.get_raw_tables()] == m['table_sha256']
    # the loaded storage works: a goetia-filled storage queries the same
    g, ref = _fill_like_goetia(m)
    sh = lg.CanLemireShifter if m['canonical'] else lg.FwdLemireShifter
    h = lg.dBG[cls, sh].build(store, m['K'])
    path = os.path.join(GOLDEN_DIR, 'data', m['input'])
    reads = read_fasta(path) if m['input'].endswith('.fa') else read_fastq(path)
    np.testing.assert_array_equal(h.query_sequences(reads), g.query_sequences(reads))
