"""Writes issue7.npz: the data of MP-sort's Issue7 test as a fixture.

Source: the Issue7B64 literal of mpsort/tests/test_mpsort.py:354-530 in the
reference tree (base64 of a protocol-3 pickle: a list of 12 per-rank lists of
numpy u4 scalars, 10 per row).  The pickle is NOT loaded: its opcode stream is
walked with pickletools.genops, only list structure is followed, and every
scalar is read from its 4-byte SHORT_BINBYTES payload; no global is looked up
and nothing from the stream is called.

The records and the expected answer are built as test_issue7 builds them
(test_mpsort.py:531-556): radix[:, 0] = A[:, 5] + (A[:, 4] << 32) in u4
arithmetic (the shift of a u4 by 32 gives 0 in numpy), radix[:, 1] = A[:, 0];
the answer is the records in numpy.lexsort order of the radix words (the last
word most significant, stable).

Run in the build container: python tests/golden/make_issue7.py
"""
import base64
import os
import pickletools
import re
import sys

import numpy

REF = os.environ.get("MPSORT_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))


class Mark(object):
    pass


def walk(raw):
    """List structure of the pickle with ('scalar', bytes) leaves."""
    stack, memo = [], {}

    def pop_mark():
        items = []
        while not isinstance(stack[-1], Mark):
            items.append(stack.pop())
        stack.pop()
        return items[::-1]

    for op, arg, _ in pickletools.genops(raw):
        name = op.name
        if name == "PROTO" or name == "STOP":
            continue
        if name == "EMPTY_LIST":
            stack.append([])
        elif name in ("BINPUT", "LONG_BINPUT"):
            memo[arg] = stack[-1]
        elif name in ("BINGET", "LONG_BINGET"):
            stack.append(memo[arg])
        elif name == "MARK":
            stack.append(Mark())
        elif name == "GLOBAL":
            stack.append(("global", arg))
        elif name in ("BINUNICODE", "SHORT_BINBYTES", "BININT1", "BININT", "NONE"):
            stack.append(arg)
        elif name == "TUPLE":
            stack.append(tuple(pop_mark()))
        elif name == "TUPLE2":
            b, a = stack.pop(), stack.pop()
            stack.append((a, b))
        elif name == "TUPLE3":
            c, b, a = stack.pop(), stack.pop(), stack.pop()
            stack.append((a, b, c))
        elif name == "REDUCE":
            args, fn = stack.pop(), stack.pop()
            stack.append(("reduce", fn, args))
        elif name == "BUILD":
            # BUILD sets the state of the object in place: memo entries
            # that refer to it see the built object
            state = stack.pop()
            obj = stack.pop()
            built = ("build", obj, state)
            for k, v in memo.items():
                if v is obj:
                    memo[k] = built
            stack.append(built)
        elif name == "APPENDS":
            items = pop_mark()
            stack[-1].extend(items)
        elif name == "APPEND":
            item = stack.pop()
            stack[-1].append(item)
        else:
            raise ValueError("unexpected opcode %s" % name)
    assert len(stack) == 1
    return stack[0]


def u4_of(leaf):
    """The value of a numpy.core.multiarray.scalar(dtype('u4'), <4 bytes>) leaf."""
    kind, fn, args = leaf
    assert kind == "reduce" and fn == ("global", "numpy.core.multiarray scalar"), leaf
    dt, payload = args
    # dt = BUILD(dtype('u4', 0, 1), state with byte order '<')
    assert dt[0] == "build" and dt[1][1] == ("global", "numpy dtype") and dt[1][2][0] == "u4"
    assert dt[2][1] == "<" and isinstance(payload, bytes) and len(payload) == 4
    return int.from_bytes(payload, "little")


def main():
    src = open(os.path.join(REF, "mpsort", "tests", "test_mpsort.py")).read()
    m = re.search(r'Issue7B64 = b"""(.*?)"""', src, re.S)
    raw = base64.decodebytes(m.group(1).encode())
    tree = walk(raw)
    assert isinstance(tree, list) and len(tree) == 12
    chunks = [numpy.array([u4_of(x) for x in rank], dtype='u4').reshape(-1, 10) for rank in tree]
    dtype = numpy.dtype([('radix', ('u8', 2)), ('ext', ('u8', 3))])
    recs = []
    for A in chunks:
        s = numpy.zeros(len(A), dtype=dtype)
        s['radix'][:, 0] = (A[:, 5]) + (A[:, 4] << 32)
        s['radix'][:, 1] = A[:, 0]
        recs.append(s)
    S = numpy.concatenate(recs)
    expected = S[numpy.lexsort(S['radix'].T)]
    counts = numpy.array([len(A) for A in chunks], dtype=numpy.int64)
    words = numpy.concatenate(chunks).astype(numpy.uint32)
    out = os.path.join(HERE, "issue7.npz")
    numpy.savez(out, counts=counts, words=words, records=S, expected=expected)
    print("wrote %s: %d ranks, counts %s, %d records of %d bytes" % (out, len(chunks), counts.tolist(),
                                                                    len(S), dtype.itemsize))


if __name__ == "__main__":
    sys.exit(main())
