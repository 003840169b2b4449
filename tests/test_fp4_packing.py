"""The fp4 operand expansion of the row search (libbicos_amd/csrc/search_mfma.hpp,
expand_fp4 / expand_c_fp4 / expand_q_fp4), checked on the CPU.

The kernel turns a 32-bit descriptor word into four registers of eight fp4
(E2M1) codes with v_perm_b32: the selector bytes of register r are the word's
bit pairs (w >> 2r) & 0x03030303, which pick bytes of a 4-byte table. The
tables are read from the header here (not retyped), v_perm_b32 is emulated
from its ISA definition (selector bytes 0..3 pick bytes 0..3 of the second
source), and the E2M1 codes are decoded from the format (sign, 2 exponent
bits with bias 1, 1 mantissa bit). What must hold for the matrix product to
rank candidates like the Hamming cost (include/impl/cpu/bicos.hpp:29-48):
candidate bits become 1.0 / 0, query bits +1.0 / -1.0, each bit of a word at
the same K position in both operands, so that the product over a word is
popcount(c) - 2 popcount(c & q)."""
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "..", "libbicos_amd", "csrc", "search_mfma.hpp")


def tables():
    src = open(HEADER).read()
    c = re.search(r"expand_c_fp4\(uint32_t w\)\s*\{\s*return expand_fp4<(0x[0-9A-Fa-f]+)u>", src)
    q = re.search(r"expand_q_fp4\(uint32_t w\)\s*\{\s*const uint4 v = expand_fp4<(0x[0-9A-Fa-f]+)u>", src)
    assert c and q, "table constants not found in the header"
    return int(c.group(1), 16), int(q.group(1), 16)


def perm(s0, s1, sel):
    """v_perm_b32 D = perm(S0, S1, S2) for selector bytes 0..7 ({S0, S1}: S1 is bytes 0..3)."""
    data = (s0 << 32) | s1
    out = 0
    for i in range(4):
        b = (sel >> (8 * i)) & 0xFF
        assert b < 8, "the expansion only uses selectors 0..3"
        out |= ((data >> (8 * b)) & 0xFF) << (8 * i)
    return out


def expand(table, w):
    m = 0x03030303
    return [perm(table, table, (w >> (2 * r)) & m) for r in range(4)]


def e2m1(code):
    """value of a 4-bit E2M1 code (OCP microscaling fp4)"""
    s = -1.0 if code & 8 else 1.0
    e, mbit = (code >> 1) & 3, code & 1
    if e == 0:
        return s * 0.5 * mbit
    return s * 2.0 ** (e - 1) * (1.0 + 0.5 * mbit)


def values(regs):
    """the 32 fp4 values of a lane's four registers, in K order (nibble j of
    register r at 8 r + j)"""
    return np.array([e2m1((regs[r] >> (4 * j)) & 0xF) for r in range(4) for j in range(8)])


def test_table_entries_encode_the_operands():
    tc, tq = tables()
    for v in range(4):  # bit pair value: low bit to the low nibble
        lo, hi = v & 1, v >> 1
        bc = (tc >> (8 * v)) & 0xFF
        bq = (tq >> (8 * v)) & 0xFF
        assert e2m1(bc & 0xF) == float(lo) and e2m1(bc >> 4) == float(hi)
        assert e2m1(bq & 0xF) == (-1.0 if lo else 1.0) and e2m1(bq >> 4) == (-1.0 if hi else 1.0)


def test_each_bit_lands_at_one_position_in_both_operands():
    tc, tq = tables()
    pos_c, pos_q = [], []
    for b in range(32):
        vc = values(expand(tc, 1 << b))
        vq = values(expand(tq, 1 << b))
        assert np.count_nonzero(vc) == 1 and np.count_nonzero(vq == -1.0) == 1
        pos_c.append(int(np.flatnonzero(vc)[0]))
        pos_q.append(int(np.flatnonzero(vq == -1.0)[0]))
    assert pos_c == pos_q and sorted(pos_c) == list(range(32))


def test_dot_is_popcount_identity():
    tc, tq = tables()
    rng = np.random.default_rng(2024)
    words = [0, 0xFFFFFFFF, 0xAAAAAAAA, 0x55555555] + [int(x) for x in rng.integers(0, 1 << 32, 400, dtype=np.uint64)]
    for i, c in enumerate(words):
        q = words[(7 * i + 3) % len(words)]
        dot = float(values(expand(tc, c)) @ values(expand(tq, q)))
        assert dot == bin(c).count("1") - 2 * bin(c & q).count("1"), (hex(c), hex(q))
