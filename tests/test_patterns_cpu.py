"""Pattern indexers: the oracle pinned by the reference's PatternTest invariants,
closed-form index-space sizes (Burnside orbit counts for VH/DC/VHDC) and BMI2
hardware PEXT/PDEP; the product's host index programs (the ones the GPU runs)
checked bit-exactly against the oracle."""
import ctypes as C
import itertools

import numpy as np
import pytest

import cassandra_amd.capi as capi
from cassandra_amd import names as N

CLASS = {"A": 0, "V": 1, "D": 2, "VH": 3, "DC": 4, "VHDC": 5}

# (name, class, index_space_size, order) from the pattern catalogue
CATALOGUE = [
    ("L0", "V", 3321, 4), ("L1", "V", 3321, 4), ("L2", "V", 3321, 4), ("L3", "V", 3321, 4),
    ("D4", "D", 45, 4), ("D5", "D", 135, 4), ("D6", "D", 378, 4), ("D7", "D", 1134, 4),
    ("D8", "DC", 3321, 2), ("L02X", "V", 29646, 4), ("B4", "A", 6561, 8), ("B5", "A", 59049, 8),
    ("B6", "A", 531441, 8), ("Q3", "D", 10206, 4), ("C3", "D", 405, 4), ("C3p1", "D", 3402, 4),
    ("C3p2", "D", 29889, 4), ("Ep", "V", 29646, 4), ("Comet", "D", 39366, 4),
]


def host_group(masks):
    arr = np.ascontiguousarray(np.array(masks, dtype=np.uint64))
    h = C.c_void_p()
    capi.call("cad_patterns_create", None, arr.ctypes.data_as(C.c_void_p), len(arr), C.byref(h))
    k, cols, npat = C.c_int(), C.c_int64(), C.c_int()
    capi.call("cad_patterns_info", h, C.byref(k), C.byref(cols), C.byref(npat))

    class G:
        pass

    g = G()
    g.h, g.k, g.cols = h, k.value, cols.value

    def indices(PO):
        PO = np.ascontiguousarray(np.asarray(PO, dtype=np.uint64).reshape(-1, 2))
        out = np.zeros((PO.shape[0], g.k), dtype=np.uint32)
        capi.call("cad_patterns_indices_host", h, PO.ctypes.data_as(C.c_void_p), PO.shape[0],
                  out.ctypes.data_as(C.c_void_p))
        return out

    def variations():
        out = np.zeros(g.k, dtype=np.uint64)
        capi.call("cad_patterns_variations", h, out.ctypes.data_as(C.c_void_p))
        return out

    def info(i):
        cls, mask, size, order, off = C.c_int(), C.c_uint64(), C.c_int64(), C.c_int(), C.c_int64()
        capi.call("cad_patterns_pattern_info", h, i, C.byref(cls), C.byref(mask), C.byref(size), C.byref(order),
                  C.byref(off))
        return dict(cls=cls.value, mask=mask.value, size=size.value, order=order.value, offset=off.value)

    g.indices, g.variations, g.info = indices, variations, info
    return g


def sym_variants(oracle, PO):
    """The 8 images of positions under D4 (PatternTest/Pattern.h:95-107)."""
    L = oracle.lib()
    V, H, D, Cd = L.or_flip_vertical, L.or_flip_horizontal, L.or_flip_diagonal, L.or_flip_codiagonal
    fs = [lambda b: b, Cd, D, H, V, lambda b: Cd(H(b)), lambda b: D(H(b)), lambda b: V(H(b))]
    return [np.array([[f(int(p)), f(int(o))] for p, o in PO], dtype=np.uint64) for f in fs]


def burnside(oracle, mask, cls):
    """Number of orbits of the 3^n configurations under the pattern's group."""
    L = oracle.lib()
    V, H, D, Cd = L.or_flip_vertical, L.or_flip_horizontal, L.or_flip_diagonal, L.or_flip_codiagonal
    if cls == "VH":
        group = [lambda b: b, V, H, lambda b: H(V(b))]
    elif cls == "DC":
        group = [lambda b: b, D, Cd, lambda b: Cd(D(b))]
    else:
        group = [lambda b: b, Cd, D, H, V, lambda b: Cd(H(b)), lambda b: D(H(b)), lambda b: V(H(b))]
    squares = [s for s in range(64) if (mask >> s) & 1]
    total = 0
    for g in group:
        # cycles of the square permutation restricted to the mask
        seen, cycles = set(), 0
        for s in squares:
            if s in seen:
                continue
            cycles += 1
            t = s
            while t not in seen:
                seen.add(t)
                t = (g(1 << t)).bit_length() - 1
        total += 3 ** cycles
    assert total % len(group) == 0
    return total // len(group)


@pytest.mark.parametrize("name,cls,size,order", CATALOGUE)
def test_catalogue(oracle, name, cls, size, order):
    m = N.PATTERNS[name]
    g = oracle.Group([m])
    info = g.info(0)
    assert (info["cls"], info["size"], info["order"]) == (CLASS[cls], size, order)
    n = bin(m).count("1")
    # closed forms
    if cls == "A":
        assert size == 3 ** n
    elif cls in ("VH", "DC", "VHDC"):
        assert size == burnside(oracle, m, cls)
    pi = host_group([m]).info(0)
    assert (pi["cls"], pi["size"], pi["order"], pi["mask"]) == (info["cls"], size, order, info["mask"])


@pytest.mark.parametrize("masks", [N.test_vh, N.test_dc, N.test_vhdc], ids=["vh", "dc", "vhdc"])
def test_dense_classes_burnside(oracle, masks):
    g = oracle.Group([masks])
    cls = {3: "VH", 4: "DC", 5: "VHDC"}[g.info(0)["cls"]]
    assert g.cols == burnside(oracle, masks, cls)


SETS = {"fit4": N.FIT4, "fit8": N.FIT8, "fit12": N.FIT12, "edax": N.edax, "logistello": N.logistello,
        "cassandra": N.cassandra, "test_mix": N.test_mix}


@pytest.mark.parametrize("name,k,cols", [("fit4", 18, 102222), ("fit8", 34, 139644), ("fit12", 54, 150282),
                                         ("edax", 46, 114363), ("logistello", 46, 113877),
                                         ("cassandra", 52, 179811)])
def test_group_shapes(oracle, name, k, cols):
    g = oracle.Group(SETS[name])
    assert (g.k, g.cols) == (k, cols)
    hg = host_group(SETS[name])
    assert (hg.k, hg.cols) == (k, cols)
    assert np.array_equal(hg.variations(), g.variations())


@pytest.mark.parametrize("name", list(SETS))
def test_product_programs_match_oracle(oracle, name):
    masks = SETS[name]
    pos = np.concatenate([oracle.random_positions_mt(1337, 2000), oracle.random_play(21, 2000, 5, 58)])
    assert np.array_equal(host_group(masks).indices(pos), oracle.Group(masks).indices(pos))


@pytest.mark.parametrize("mask", N.test_mix + [N.test_mix], ids=["a", "c", "d", "h", "v", "vh", "dc", "vhdc", "mix"])
def test_symmetry_independence(oracle, mask):
    # PatternTest/Indexer.cpp:8-47 (RandomPositionGenerator(1337), 10,000 positions)
    masks = mask if isinstance(mask, list) else [mask]
    pos = oracle.random_positions_mt(1337, 10000)
    for idx_of in (oracle.Group(masks).indices, host_group(masks).indices):
        base = np.sort(idx_of(pos), axis=1)
        for var in sym_variants(oracle, pos):
            assert np.array_equal(np.sort(idx_of(var), axis=1), base)


@pytest.mark.parametrize("mask", N.test_mix + [N.test_mix], ids=["a", "c", "d", "h", "v", "vh", "dc", "vhdc", "mix"])
def test_configurations_cover_index_space(oracle, mask):
    # PatternTest/Indexer.cpp:51-92
    masks = mask if isinstance(mask, list) else [mask]
    union = 0
    for m in masks:
        union |= m
    cfg = oracle.configurations(union)
    for g in (oracle.Group(masks), host_group(masks)):
        seen = np.unique(g.indices(cfg))
        assert np.array_equal(seen, np.arange(g.cols))


def test_configurations_order(oracle):
    # Pattern/Helpers.cpp:81-102: p outer, o inner, first item (0, 0)
    cfg = oracle.configurations(0b101)
    want = []
    for p in range(4):
        for o in range(4):
            if p & o == 0:
                want.append((oracle.lib().or_pdep(p, 0b101), oracle.lib().or_pdep(o, 0b101)))
    assert [tuple(map(int, c)) for c in cfg] == want
    assert len(oracle.configurations(N.L02X)) == 3 ** 10


def test_fast_index_closed_form(oracle):
    L = oracle.lib()
    for m in [N.test_a, N.D4, N.C3, N.Q3, 0x8000000000000001, N.L02X]:
        squares = [s for s in range(64) if (m >> s) & 1]
        for p, o in oracle.configurations(m)[::7]:
            p, o = int(p), int(o)
            want = sum(3 ** i * (((p >> s) & 1) + 2 * ((o >> s) & 1)) for i, s in enumerate(squares))
            assert L.or_fast_index(p, o, m) == want


def test_pext_pdep_against_bmi2(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(5)
    a = rng.integers(0, 2**63, size=(20000, 2), dtype=np.int64).astype(np.uint64) * np.uint64(2) + np.uint64(1)
    for s, m in a:
        s = int(s)
        m = int(m) & ((s * 7 + 3) & ((1 << 64) - 1))
        assert L.or_pext(s, m) == L.or_hw_pext(s, m)
        assert L.or_pdep(s, m) == L.or_hw_pdep(s, m)


def test_pattern_errors():
    # more than 16 squares cannot be indexed with a 2^16 power-sum table
    arr = np.array([0xFFFFFFFF], dtype=np.uint64)
    h = C.c_void_p()
    rc = capi.lib().cad_patterns_create(None, arr.ctypes.data_as(C.c_void_p), 1, C.byref(h))
    assert rc == -3 and b"16" in capi.lib().cad_last_error()
    rc = capi.lib().cad_patterns_create(None, arr.ctypes.data_as(C.c_void_p), 0, C.byref(h))
    assert rc == -3


def test_evaluator_identities(oracle):
    # PatternTest/Evaluator.cpp:7-31: iota weights round-trip through the
    # expanded tables; Eval is symmetry-invariant; Eval(pos) == (A x)[row]
    # bit-exactly in FP32 (same terms, same order).
    for mask in N.test_mix:
        g = oracle.Group([mask])
        w = np.arange(1, g.cols + 1, dtype=np.float32)
        tables = g.expand(w)
        assert np.array_equal(g.compact(tables), w)
        cfg = oracle.configurations(mask)
        base = g.eval_expanded(tables, cfg)
        for var in sym_variants(oracle, cfg):
            assert np.array_equal(g.eval_expanded(tables, var), base)
    g = oracle.Group(N.FIT4)
    rng = np.random.default_rng(3)
    w = rng.normal(size=g.cols).astype(np.float32)
    pos = oracle.random_play(9, 3000, 20)
    y_eval = g.eval_expanded(g.expand(w), pos)
    y_spmv = oracle.spmv(g.indices(pos), w)
    assert np.array_equal(y_eval, y_spmv)


# ---- exact numbering, pinned by order: each class's index is the rank of a
# canonical key in a stated order, which a permuted or transposed formula breaks.

def _square_map(f):
    """Permutation of bit numbers b = x + 8y under a coordinate map (x, y) -> f(x, y)."""
    out = []
    for b in range(64):
        x, y = f(b % 8, b // 8)
        out.append(x + 8 * y)
    return out


_MAPS = {
    "V": _square_map(lambda x, y: (x, 7 - y)),
    "H": _square_map(lambda x, y: (7 - x, y)),
    "D": _square_map(lambda x, y: (y, x)),
    "C": _square_map(lambda x, y: (7 - y, 7 - x)),
}


def _flip(a, t):
    a = np.asarray(a, dtype=np.uint64)
    out = np.zeros_like(a)
    for s, d in enumerate(_MAPS[t]):
        out |= ((a >> np.uint64(s)) & np.uint64(1)) << np.uint64(d)
    return out


def _apply(PO, ts):
    """T1(T2(...(pos))) for ts = 'T1T2...' applied right to left, to P and O."""
    P, O = PO[:, 0].copy(), PO[:, 1].copy()
    for t in reversed(ts):
        P, O = _flip(P, t), _flip(O, t)
    return np.stack([P, O], axis=1)


def _fast_index(PO, mask):
    squares = [s for s in range(64) if (mask >> s) & 1]
    idx = np.zeros(PO.shape[0], dtype=np.int64)
    for i, s in enumerate(squares):
        p = ((PO[:, 0] >> np.uint64(s)) & np.uint64(1)).astype(np.int64)
        o = ((PO[:, 1] >> np.uint64(s)) & np.uint64(1)).astype(np.int64)
        idx += 3 ** i * (p + 2 * o)
    return idx


def _lex_rank(keys):
    """Rank of each row of keys among all rows of the full key space, which is
    given as the distinct rows of keys themselves (the configurations cover it)."""
    uniq, inv = np.unique(keys, axis=0, return_inverse=True)  # np.unique sorts rows lexicographically
    return inv.reshape(-1), len(uniq)


# InsertIndices orders (Pattern/Indexer.cpp:31-41, 82-88, 128-134 and the dense
# classes' lists): column j of a pattern's block is variation 0 of T_j(pos).
_ORDERS = {0: ["", "C", "D", "H", "V", "HC", "HD", "HV"], 1: ["", "C", "D", "H"], 2: ["", "C", "H", "V"],
           3: ["", "D"], 4: ["", "V"], 5: [""]}


@pytest.mark.parametrize("mask", N.test_mix + [N.L02X, N.Q3, N.D8, N.B4, N.C3p1],
                         ids=["a", "c", "d", "h", "v", "vh", "dc", "vhdc", "L02X", "Q3", "D8", "B4", "C3p1"])
def test_index_numbering_is_canonical(oracle, mask):
    g = oracle.Group([mask])
    info = g.info(0)
    cls, p = info["cls"], info["mask"]
    # dispatch (Pattern/Indexer.cpp:345-364): h-only indexes D(p) as V, c-only V(p) as D
    sym = {t: int(_flip(np.array([mask], np.uint64), t)[0]) == mask for t in "VHDC"}
    if sym["H"] and not sym["V"]:
        assert p == int(_flip(np.array([mask], np.uint64), "D")[0])
    elif sym["C"] and not sym["D"]:
        assert p == int(_flip(np.array([mask], np.uint64), "V")[0])
    else:
        assert p == mask
    cfg = oracle.configurations(p).astype(np.uint64)
    got = [host_group([mask]).indices(cfg), g.indices(cfg)]
    assert np.array_equal(got[0], got[1])
    idx0 = got[0][:, 0].astype(np.int64)
    if cls == 0:
        want = _fast_index(cfg, p)
    elif cls == 1:  # V: rank of the unordered pair {lower half, lower half of V(pos)}
        half = p & 0x00000000FFFFFFFF
        a, b = _fast_index(cfg, half), _fast_index(_apply(cfg, "V"), half)
        want, n = _lex_rank(np.stack([np.minimum(a, b), np.maximum(a, b)], axis=1))
        hs = 3 ** bin(half).count("1")
        assert n == hs * (hs + 1) // 2
    elif cls == 2:  # D: rank of (unordered pair of triangle halves, diagonal)
        half, diag = p & 0x0080C0E0F0F8FCFE, p & 0x8040201008040201
        a, b = _fast_index(cfg, half), _fast_index(_apply(cfg, "D"), half)
        want, n = _lex_rank(np.stack([np.minimum(a, b), np.maximum(a, b), _fast_index(cfg, diag)], axis=1))
        hs, ds = 3 ** bin(half).count("1"), 3 ** bin(diag).count("1")
        assert n == ds * hs * (hs + 1) // 2
    else:  # dense: orbit ids handed out in first-seen Configurations order
        seen, nxt, want = {}, 0, np.empty(len(cfg), np.int64)
        group = {3: ["", "V", "H", "HV"], 4: ["", "D", "C", "CD"],
                 5: ["", "C", "D", "H", "V", "CH", "DH", "VH"]}[cls]
        images = np.stack([_fast_index(_apply(cfg, t), p) for t in group], axis=1)
        for i, f in enumerate(_fast_index(cfg, p)):
            if f not in seen:
                for im in images[i]:
                    seen[int(im)] = nxt
                nxt += 1
            want[i] = seen[f]
    assert np.array_equal(idx0, want)
    # the other columns are variation 0 of the transformed position
    pos = np.concatenate([oracle.random_positions_mt(14, 3000), oracle.random_play(127, 2000, 10, 50)])
    rows = host_group([mask]).indices(pos)
    for j, t in enumerate(_ORDERS[cls]):
        assert np.array_equal(rows[:, j], host_group([mask]).indices(_apply(pos, t))[:, 0])
