"""Pins of the oracle's RNG (numba.cuda.random xoroshiro128+, splitmix64 seeding,
2**64 jump, float conversions) against published values and mathematics.

No reference test pins RNG values (only tests/test_parallel_tempering.py:70,80-81
"exchange leaves site RNG untouched"), and numba itself cannot be imported
here, so the pins are:
* splitmix64 known answers for state 0 (Vigna's splitmix64.c output stream);
* next() against an independent big-integer restatement;
* jump() equals T**(2**64) for the transition matrix T over GF(2), computed
  by 64 matrix squarings (independent of the published jump polynomial);
* float conversions against numpy's round-to-nearest casts, including the
  float32 rounding of the largest draws to 1.0.
"""

import numpy as np
import pytest

import oracle

MASK = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15


def test_splitmix64_known_answers():
    # First outputs of splitmix64 from state 0: the generator adds GOLDEN
    # before mixing, so output k is the mix of k*GOLDEN + GOLDEN.
    expected = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC]
    got = [oracle.splitmix64((k * GOLDEN) & MASK) for k in range(4)]
    assert [hex(x) for x in got] == [hex(x) for x in expected]


def _rotl(x, k):
    return ((x << k) | (x >> (64 - k))) & MASK


def _py_next(s0, s1):
    result = (s0 + s1) & MASK
    s1 ^= s0
    return result, _rotl(s0, 55) ^ s1 ^ ((s1 << 14) & MASK), _rotl(s1, 36)


def test_next_matches_big_integer_restatement():
    states = oracle.rng_init(3, 12345)
    for k in range(3):
        s0, s1 = int(states[k, 0]), int(states[k, 1])
        expected = []
        for _ in range(200):
            r, s0, s1 = _py_next(s0, s1)
            expected.append(r)
        got = oracle.rng_draw_u64(states, k, 200)
        assert [int(x) for x in got] == expected
        assert (int(states[k, 0]), int(states[k, 1])) == (s0, s1)


def _apply(cols, v):
    r = 0
    b = 0
    while v:
        if v & 1:
            r ^= cols[b]
        v >>= 1
        b += 1
    return r


def _transition_columns():
    cols = []
    for b in range(128):
        v = 1 << b
        s0, s1 = v & MASK, v >> 64
        _, n0, n1 = _py_next(s0, s1)
        cols.append(n0 | (n1 << 64))
    return cols


def test_jump_equals_two_to_the_64_steps_over_gf2():
    cols = _transition_columns()
    for _ in range(64):  # T -> T**2, 64 times: T**(2**64)
        cols = [_apply(cols, c) for c in cols]
    rng = np.random.default_rng(3)
    for _ in range(8):
        s = rng.integers(0, 2**64 - 1, size=2, dtype=np.uint64, endpoint=True)
        v = int(s[0]) | (int(s[1]) << 64)
        expected = _apply(cols, v)
        jumped = oracle.rng_jump(s)
        assert int(jumped[0]) | (int(jumped[1]) << 64) == expected


def test_init_layout_seed_wrap_and_subsequences():
    for seed in (0, 42, -1, 2**63 + 7):
        states = oracle.rng_init(5, seed)
        z = oracle.splitmix64(seed & MASK)
        assert int(states[0, 0]) == z and int(states[0, 1]) == z
        for k in range(1, 5):
            np.testing.assert_array_equal(states[k], oracle.rng_jump(states[k - 1]))
    np.testing.assert_array_equal(oracle.rng_init(4, 9, subsequence_start=2)[0],
                                  oracle.rng_init(3, 9)[2])
    assert oracle.rng_init(0, 1).shape == (0, 2)


def test_uniform_conversions_round_to_nearest():
    states = oracle.rng_init(1, 77)
    raw = oracle.rng_draw_u64(states.copy(), 0, 5000)
    f64 = oracle.rng_draw_f64(states.copy(), 0, 5000)
    f32 = oracle.rng_draw_f32(states.copy(), 0, 5000)
    expected64 = (raw >> np.uint64(11)).astype(np.float64) * 2.0**-53
    np.testing.assert_array_equal(f64, expected64)
    np.testing.assert_array_equal(f32, expected64.astype(np.float32))
    assert f64.min() >= 0 and f64.max() < 1
    # float32 rounding sends (2**53 - 1) * 2**-53 to exactly 1.0
    top = np.float64((2**53 - 1) * 2.0**-53)
    assert np.float32(top) == np.float32(1.0)
