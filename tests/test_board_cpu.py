"""Board primitives: the oracle pinned by the reference's BoardTest known answers,
and the product's host board code (the same source the GPU kernels compile)
checked against the oracle."""
import numpy as np
import pytest

import cassandra_amd.capi as capi

START_P, START_O = 0x0000000810000000, 0x0000001008000000
FLIP = {"vertical": 1, "horizontal": 2, "diagonal": 3, "codiagonal": 4}


def test_flip_known_answers(oracle):
    # BoardTest/BitBoard.cpp:8-23
    L = oracle.lib()
    kat = {"codiagonal": 0x8080808000000000, "diagonal": 0x01010101, "horizontal": 0xF0,
           "vertical": 0x0F00000000000000}
    for name, want in kat.items():
        assert getattr(L, f"or_flip_{name}")(0xF) == want
        assert capi.lib().cad_flip(FLIP[name], 0xF) == want


def test_flips_single_squares_and_group_laws(oracle):
    L = oracle.lib()
    maps = {"vertical": lambda x, y: (x, 7 - y), "horizontal": lambda x, y: (7 - x, y),
            "diagonal": lambda x, y: (y, x), "codiagonal": lambda x, y: (7 - y, 7 - x)}
    for name, f in maps.items():
        for s in range(64):
            nx, ny = f(s % 8, s // 8)
            want = 1 << (nx + 8 * ny)
            assert getattr(L, f"or_flip_{name}")(1 << s) == want
            assert capi.lib().cad_flip(FLIP[name], 1 << s) == want
    rng = np.random.default_rng(0)
    for b in rng.integers(0, 2**63, size=200, dtype=np.int64).astype(np.uint64):
        b = int(b) * 2 + 1
        b &= (1 << 64) - 1
        F = {n: (lambda v, n=n: capi.lib().cad_flip(FLIP[n], v)) for n in FLIP}
        for n in FLIP:  # involutions
            assert F[n](F[n](b)) == b
        # D then C is the 180 degree turn, as is H then V
        assert F["codiagonal"](F["diagonal"](b)) == F["vertical"](F["horizontal"](b))


def test_possible_moves_start_position(oracle):
    # BoardTest/PossibleMoves.cpp:56
    assert oracle.lib().or_possible_moves(START_P, START_O) == 0x0000102004080000
    assert capi.lib().cad_possible_moves(START_P, START_O) == 0x0000102004080000


def test_random_position_generator_uses_raw_engine_output(oracle):
    # Board/PositionGenerator.h:9-10: uniform_int_distribution over the full
    # 64-bit range returns the engine's draws unchanged (libstdc++)
    a = oracle.random_positions_mt(1337, 1000)
    b = oracle.random_positions_mt_dist(1337, 1000)
    assert np.array_equal(a, b)
    assert np.all((a[:, 0] & a[:, 1]) == 0)


def test_flips_against_loop_reference(oracle):
    # BoardTest/Flips.cpp:38-48: 10k positions of RandomPositionGenerator(14)
    # per move; the product's flips against the loop form.
    pos = oracle.random_positions_mt(14, 2000)
    L, P = oracle.lib(), capi.lib()
    for (p, o) in pos:
        p, o = int(p), int(o)
        empty = ~(p | o) & ((1 << 64) - 1)
        for m in range(64):
            if (empty >> m) & 1:
                assert P.cad_flips(p, o, m) == L.or_flips(p, o, m)


def test_possible_moves_random_samples(oracle):
    # BoardTest/PossibleMoves.cpp:61-74 (seed 127): moves == squares with flips
    pos = oracle.random_positions_mt(127, 5000)
    L, P = oracle.lib(), capi.lib()
    for (p, o) in pos:
        p, o = int(p), int(o)
        assert P.cad_possible_moves(p, o) == L.or_possible_moves(p, o)


@pytest.mark.parametrize("plies,count", [(1, 4), (2, 12), (3, 56), (4, 244), (5, 1396), (6, 8200), (7, 55092)])
def test_perft_counts(oracle, plies, count):
    # BoardTest/Children.cpp:10-17 (NumberOfChildren e59..e53)
    assert oracle.lib().or_perft(plies) == count


def test_random_play_host_matches_oracle(oracle):
    import ctypes as C

    for seed, n, lo, hi in [(1, 500, 24, 24), (7, 500, 10, 59), (3, 200, 60, 60), (11, 300, 15, 45)]:
        out = np.zeros((n, 2), np.uint64)
        capi.call("cad_random_play_host", seed, n, lo, hi, out.ctypes.data_as(C.c_void_p))
        ref = oracle.random_play(seed, n, lo, hi)
        assert np.array_equal(out, ref)
        empties = [bin(~(int(p) | int(o)) & ((1 << 64) - 1)).count("1") for p, o in out]
        assert min(empties) >= lo and max(empties) <= hi
        assert np.all((out[:, 0] & out[:, 1]) == 0)
