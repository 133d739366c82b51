"""CPU restatement of goetia's FASTX parsing -- TEST INFRASTRUCTURE ONLY.

What a caller looping on FastxParser<DNA_SIMPLE>::next (readers.hh:148-219)
receives, restated byte by byte from the vendored kseq.h reader it wraps:

- kstream (kseq.h:62-140): ks_getc, and ks_getuntil2 with the space and line
  delimiters; a line's trailing '\\r' is dropped when the accumulated string is
  longer than one byte (kseq.h:137).
- kseq_read (kseq.h:173-211): skip to a '>' or '@'; name up to white space,
  comment to the end of the line; sequence lines until a line starting with
  '>', '+' or '@' (empty lines skipped); FASTQ: skip the '+' line, read quality
  lines until the quality is at least as long as the sequence; -2 when it is
  missing or of another length.
- FastxParser::next: DNA_SIMPLE::validate (alphabets.hh:41-55, 112-130)
  uppercases a/c/g/t and rejects other bytes (n_skipped += 1; strict: stop);
  records shorter than min_length are skipped (n_skipped += 1, also on top of
  a rejection); n_parsed counts records read with kseq status >= 0 that did
  not stop a strict parse; an InvalidRead (-2) counts as skipped.

Pinned against the compiled reference's parser output on the fixtures in
tests/golden/data (tests/golden/fastx_ref.json).  Only tests import this.
"""
import gzip

SPACE = b' \t\n\v\f\r'
VALID = {ord('A'): 'A', ord('C'): 'C', ord('G'): 'G', ord('T'): 'T',
         ord('a'): 'A', ord('c'): 'C', ord('g'): 'G', ord('t'): 'T'}


class _Stream:
    def __init__(self, data):
        self.d = data
        self.i = 0

    def getc(self):
        if self.i >= len(self.d):
            return -1
        c = self.d[self.i]
        self.i += 1
        return c

    def getuntil(self, line, s, append):
        """Returns (length or -1, delimiter byte or 0, new s)."""
        if not append:
            s = bytearray()
        if self.i >= len(self.d):
            return -1, 0, s
        j = self.i
        if line:
            k = self.d.find(b'\n', j)
            end = len(self.d) if k < 0 else k
        else:
            end = j
            while end < len(self.d) and self.d[end] not in SPACE:
                end += 1
        s += self.d[j:end]
        dret = self.d[end] if end < len(self.d) else 0
        self.i = end + 1
        if line and len(s) > 1 and s[-1] == ord('\r'):
            s = s[:-1]
        return len(s), dret, s


def parse(path, strict=False, min_length=0, with_quality=False):
    data = open(path, 'rb').read()
    if data[:2] == b'\x1f\x8b':
        data = gzip.decompress(data)
    ks = _Stream(data)
    last_char = 0
    out = {'sequences': [], 'names': [], 'n_parsed': 0, 'n_skipped': 0, 'invalid_reads': 0, 'strict_stopped': 0}
    if with_quality:
        out['qualities'] = []     # Record::quality (readers.hh:186-191); None when empty
    while True:
        # kseq_read
        stat = None
        if last_char == 0:
            c = ks.getc()
            while c >= 0 and c not in (ord('>'), ord('@')):
                c = ks.getc()
            if c < 0:
                stat = c
            else:
                last_char = c
        if stat is None:
            r, c, name = ks.getuntil(False, bytearray(), False)
            if r < 0:
                stat = r
        if stat is None:
            if c != ord('\n'):
                ks.getuntil(True, bytearray(), False)
            seq = bytearray()
            c = ks.getc()
            while c >= 0 and c not in (ord('>'), ord('+'), ord('@')):
                if c != ord('\n'):
                    seq.append(c)
                    _, _, seq = ks.getuntil(True, seq, True)
                c = ks.getc()
            if c in (ord('>'), ord('@')):
                last_char = c
            qual = bytearray()
            if c != ord('+'):
                stat = len(seq)
            else:
                c = ks.getc()
                while c >= 0 and c != ord('\n'):
                    c = ks.getc()
                if c == -1:
                    stat = -2
                else:
                    qual = bytearray()
                    while True:
                        r, _, qual = ks.getuntil(True, qual, True)
                        if not (r >= 0 and len(qual) < len(seq)):
                            break
                    last_char = 0
                    stat = -2 if len(seq) != len(qual) else len(seq)
        # FastxParser::next
        if stat >= 0:
            bad = any(b not in VALID for b in seq)
            if bad:
                out['n_skipped'] += 1
                if strict:
                    out['strict_stopped'] = 1
                    return out
                stat = -4
            if len(seq) < min_length:
                stat = -4
                out['n_skipped'] += 1
            if stat >= 0:
                out['sequences'].append(''.join(VALID[b] for b in seq))
                out['names'].append(bytes(name).decode())
                if with_quality:
                    out['qualities'].append(bytes(qual).decode() if qual else None)
            out['n_parsed'] += 1
        if stat == -1:
            return out
        if stat == -2:
            out['n_skipped'] += 1
            out['invalid_reads'] += 1
