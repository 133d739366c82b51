"""Writes the FASTX parsing fixtures and their golden outputs.

Inputs (tests/golden/data/fastx/, written here, seeded): hand-made FASTA and
FASTQ files covering multi-line records, empty lines and records, CRLF line
ends, header comments, lowercase bases, N / IUPAC / space bytes, a FASTQ
record whose quality is shorter than its sequence, a quality line starting
with '@', a gzip copy, and a seeded multi-record file with random line
wrapping; plus goetia's own test files (tests/golden/data/).

Outputs (tests/golden/fastx_ref.json): for each file and (strict, min_length)
setting, what goetia's FastxParser<DNA_SIMPLE> returns record by record
(readers.hh:148-219, over the vendored kseq.h), run through the compiled
reference (oracle/_ref/libgoetia_ref.so, ref_fastx_parse in
oracle/ref_driver.cc): sequences, names, n_parsed, n_skipped, InvalidRead
count, and whether strict mode stopped on an invalid character.

    python tests/golden/make_fastx_golden.py     (needs `make -C oracle ref`)
"""
import ctypes as C
import gzip
import json
import os
import random

HERE = os.path.dirname(os.path.abspath(__file__))
DATA = os.path.join(HERE, 'data')
FX = os.path.join(DATA, 'fastx')
REF_SO = os.path.join(HERE, '..', '..', 'oracle', '_ref', 'libgoetia_ref.so')

MIXED_FA = (
    b">r1 first read, with a comment\n"
    b"ACGTACGTAC\n"
    b"GTACGTACGT\n"
    b"\n"
    b">r2\n"
    b"acgtnACGT\n"
    b">r3 lower case\n"
    b"acgtacgtacgtacgtacgtacgt\n"
    b"ACG\n"
    b">empty\n"
    b">r5\tdesc\r\n"
    b"ACGTTTGCAACGTTTGCAACGTTTGCAACGTTTGCAACGTTTGCAA\r\n"
    b"CCCCGGGGAAAATTTT\r\n"
    b">r6 space inside\n"
    b"ACGT ACGT\n"
    b">r7\n"
    b"NNNN\n"
    b">r8 iupac\n"
    b"ACGTRYACGT\n"
    b">r9\n"
    b"TTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTTT\n"
    b">r10 trailing\n"
    b"GATTACA"
)

MIXED_FQ = (
    b"@q1 comment here\n"
    b"ACGTACGTACGTACGTACGTACGTACGTACGTACGT\n"
    b"+\n"
    b"IIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIII\n"
    b"@q2\n"
    b"acgtacgtacgtacgtacgtACGTACGTACGTACGT\n"
    b"+q2\n"
    b"@IIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIII\n"
    b"@q3 with N\n"
    b"ACGTNACGTACGTACGTACGTACGTACGTACGTAC\n"
    b"+\n"
    b"IIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIIII\n"
    b"@q4 multi-line\n"
    b"ACGTACGTACGTACGTACGT\n"
    b"CCGGTTAACCGGTTAA\n"
    b"+\n"
    b"IIIIIIIIIIIIIIIIIIII\n"
    b"IIIIIIIIIIIIIIII\n"
    b"@q5 short quality\n"
    b"ACGTACGTACGTACGT\n"
    b"+\n"
    b"IIII\n"
    b"@q6 crlf\r\n"
    b"GGGGCCCCAAAATTTTGGGGCCCCAAAATTTT\r\n"
    b"+\r\n"
    b"IIIIIIIIIIIIIIIIIIIIIIIIIIIIIIII\r\n"
    b"@q7\n"
    b"ACGTACGTAC\n"
    b"+\n"
    b"IIIIIIIIII\n"
)


def random_wrapped(seed, n):
    rng = random.Random(seed)
    out = []
    for i in range(n):
        L = rng.choice([0, 5, 20, 31, 60, 150, 151, 300, 1000])
        s = ''.join(rng.choice('ACGT') for _ in range(L))
        if rng.random() < 0.2:
            s = s.lower()
        if L and rng.random() < 0.1:
            j = rng.randrange(L)
            s = s[:j] + rng.choice('NRYKM.-*') + s[j + 1:]
        header = f'>rnd{i}' + (f' len={L}' if rng.random() < 0.5 else '')
        out.append(header)
        w = rng.choice([60, 70, 80, 1000])
        for k in range(0, len(s), w):
            out.append(s[k:k + w])
        if rng.random() < 0.05:
            out.append('')
    return ('\n'.join(out) + '\n').encode()


def write_inputs():
    os.makedirs(FX, exist_ok=True)
    files = {'mixed.fa': MIXED_FA, 'mixed.fq': MIXED_FQ, 'random_wrapped.fa': random_wrapped(7, 500)}
    for name, data in files.items():
        with open(os.path.join(FX, name), 'wb') as f:
            f.write(data)
    with gzip.GzipFile(os.path.join(FX, 'mixed.fq.gz'), 'wb', mtime=0) as f:
        f.write(MIXED_FQ)
    return ['fastx/mixed.fa', 'fastx/mixed.fq', 'fastx/mixed.fq.gz', 'fastx/random_wrapped.fa',
            'random-20-a.fa', 'test-fastq-reads.fq']


def parse_ref(lib, path, strict, min_length):
    size = os.path.getsize(path) * 8 + 1024
    seq = C.create_string_buffer(size)
    names = C.create_string_buffer(size)
    cap = 1 << 16
    offsets = (C.c_int64 * (cap + 1))()
    name_off = (C.c_int64 * (cap + 1))()
    counts = (C.c_uint64 * 4)()
    n = lib.ref_fastx_parse(path.encode(), strict, min_length, seq, size, offsets, names, size, name_off, cap,
                            counts)
    assert n <= cap
    return {
        'sequences': [seq.raw[offsets[i]:offsets[i + 1]].decode() for i in range(n)],
        'names': [names.raw[name_off[i]:name_off[i + 1]].decode() for i in range(n)],
        'n_parsed': counts[0], 'n_skipped': counts[1], 'invalid_reads': counts[2], 'strict_stopped': counts[3],
    }


def main():
    files = write_inputs()
    lib = C.CDLL(REF_SO)
    lib.ref_fastx_parse.restype = C.c_uint64
    lib.ref_fastx_parse.argtypes = [C.c_char_p, C.c_int, C.c_uint32, C.c_char_p, C.c_uint64,
                                    C.POINTER(C.c_int64), C.c_char_p, C.c_uint64, C.POINTER(C.c_int64),
                                    C.c_uint64, C.POINTER(C.c_uint64)]
    out = {}
    for rel in files:
        path = os.path.join(DATA, rel)
        for strict in (0, 1):
            for min_length in (0, 31):
                out[f'{rel}|{strict}|{min_length}'] = parse_ref(lib, path, strict, min_length)
    with open(os.path.join(HERE, 'fastx_ref.json'), 'w') as f:
        json.dump(out, f, indent=0, sort_keys=True)
    print(f'{len(out)} parses written')


if __name__ == '__main__':
    main()
