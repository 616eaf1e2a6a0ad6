"""The oracle's seeded random unit vectors (the planner's RandomSeeded state,
crates/fullmag-plan/src/util.rs:118-147), pinned independently of the
restatement.  The reference's own test checks only count and norm
(crates/fullmag-plan/src/tests.rs:1927-1935), and no Rust toolchain is here
to run it, so:

- the state recurrence (util.rs:128-130) is Marsaglia's xor64 (shifts 13, 7,
  17, left-right-left; G. Marsaglia, "Xorshift RNGs", J. Stat. Softw. 8(14),
  2003).  Its published seed 88172645463325252 gives the first outputs
  8748534153485358512, 3040900993826735515, 3453997556048239312, the values
  other implementations use as test vectors; a wrong shift, direction or
  order of the three steps changes them;
- the recurrence is linear over GF(2); its 64 x 64 matrix must have the
  full multiplicative order 2^64 - 1 (Marsaglia's condition for the
  maximal period), which only a few shift triples satisfy;
- the component map (util.rs:133) at states whose values are exact:
  2^63 -> 0, 2^64 - 1 -> 1 (u64::MAX as f64 is 2^64), 2^62 -> -1/2,
  3 * 2^62 -> 1/2;
- the vectors are the normalised triples of consecutive states.
"""
import numpy as np

from oracle import oracle as O

M64 = (1 << 64) - 1


def test_marsaglia_xor64_known_outputs():
    assert O.xorshift64_states(88172645463325252, 3) == [8748534153485358512, 3040900993826735515,
                                                        3453997556048239312]


def _matmul(a, b):
    # GF(2) 64 x 64 matrices as lists of 64 row bitmasks: (a b)[i] = xor of
    # b's rows j where bit j of a[i] is set
    out = []
    for row in a:
        acc, j = 0, 0
        while row:
            if row & 1:
                acc ^= b[j]
            row >>= 1
            j += 1
        out.append(acc)
    return out


def _matpow(m, e):
    r = [1 << i for i in range(64)]
    while e:
        if e & 1:
            r = _matmul(r, m)
        m = _matmul(m, m)
        e >>= 1
    return r


def test_recurrence_has_full_period_over_gf2():
    # column i of the step's matrix T is the step applied to basis state
    # e_i; t[r] holds row r of T as a bitmask
    cols = [O.xorshift64_states(1 << i, 1)[0] for i in range(64)]
    t = [sum(((cols[i] >> r) & 1) << i for i in range(64)) for r in range(64)]
    # the step is linear: state(a ^ b) = state(a) ^ state(b)
    a, b = 0x0123456789ABCDEF, 0xF0E1D2C3B4A59687
    assert O.xorshift64_states(a ^ b, 1)[0] == O.xorshift64_states(a, 1)[0] ^ O.xorshift64_states(b, 1)[0]
    ident = [1 << i for i in range(64)]
    order = M64
    assert _matpow(t, order) == ident
    for p in (3, 5, 17, 257, 641, 65537, 6700417):   # the primes of 2^64 - 1
        assert _matpow(t, order // p) != ident, p


def test_component_map_exact_values():
    assert O.state_component(1 << 63) == 0.0
    assert O.state_component(M64) == 1.0
    assert O.state_component(1 << 62) == -0.5
    assert O.state_component(3 << 62) == 0.5
    assert O.state_component(1) == -1.0


def test_vectors_are_normalised_consecutive_state_triples():
    seed, n = 42, 50
    states = O.xorshift64_states(seed, 3 * n)
    v = O.random_unit_vectors(seed, n)
    for i in range(n):
        c = np.array([O.state_component(s) for s in states[3 * i:3 * i + 3]])
        assert np.array_equal(v[i], c / np.sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]))
