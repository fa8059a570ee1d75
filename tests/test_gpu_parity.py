"""GPU parity: every stage of the HIP server path against the CPU oracle,
bit for bit, on the same seeded keys, query and database.

Shapes: the reference's CPU-runnable configuration (fst 256 x other 32,
N = 2048, t = 65537, l = 5, l_key = 10) and a small odd shape (fst 12 x
other 4) that exercises the generic first-dimension kernel.
"""
import numpy as np
import pytest

from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

N = 2048
DB_SEED = 7
CLIENT_SEED = 11
CID = 5


@pytest.fixture(scope="module")
def cfg1():
    from onionpirv2_amd import GpuServer
    o = Oracle(fst=256, other=32)
    srv = GpuServer(256, 32)
    sk = o.keygen(CLIENT_SEED)
    gk = o.galois_keys(CLIENT_SEED, sk)
    rk = o.rgsw_key(CLIENT_SEED, sk)
    srv.set_galois_key(CID, gk)
    srv.set_rgsw_key(CID, rk)
    srv.db_generate(DB_SEED, 0)
    db = o.db_build(DB_SEED, 0)
    yield dict(o=o, srv=srv, sk=sk, gk=gk, rk=rk, rkn=o.rgsw_key_ntt(rk), db=db)
    srv.close()


def test_derived_parameters(cfg1):
    o, srv = cfg1["o"], cfg1["srv"]
    i = srv.info
    assert (i["q"], i["p"], i["t"], i["small_q"]) == (o.q, o.p, o.t, o.small_q)
    assert (i["base_log2"], i["base_log2_key"], i["ndim"], i["expan_height"], i["useful_cnt"]) == \
        (o.base_log2, o.base_log2_key, o.ndim, o.h, o.u)


@pytest.mark.parametrize("which", [0, 1])
@pytest.mark.parametrize("radix", [8, 4])
def test_ntt_matches_oracle(cfg1, which, radix):
    """Forward and inverse transforms mod q and mod p in both pass schedules
    (radix-8: 3+3+3+2 stages; radix-4: 2+2+2+2+3)."""
    o, srv = cfg1["o"], cfg1["srv"]
    rng = np.random.default_rng(1 + which)
    m = o.q if which == 0 else o.p
    polys = rng.integers(0, m, size=(8, N), dtype=np.uint64)
    polys[0] = 0
    polys[1] = m - 1
    saved = srv.get_tuning()
    try:
        srv.set_tuning(ntt_radix=radix)
        fwd = srv.tap_ntt(polys, which=which)
        exp = np.stack([o.ntt(p, which) for p in polys]).reshape(-1)
        assert np.array_equal(fwd, exp)
        inv = srv.tap_ntt(fwd, which=which, inverse=True)
        assert np.array_equal(inv, polys.reshape(-1))
    finally:
        srv.set_tuning(**saved)


def test_galois_keyswitch_matches_oracle(cfg1):
    o, srv = cfg1["o"], cfg1["srv"]
    ct = o.encrypt(CLIENT_SEED, 900, cfg1["sk"], np.arange(N, dtype=np.uint64) % o.t)
    import ctypes as C
    from oracle_lib import ptr, u64
    for level in range(o.h):
        g = N // (1 << level) + 1
        exp = u64(2 * N)
        o.L.orc_apply_galois(o.c, ptr(ct), g, ptr(cfg1["gk"][level * 4 * N:(level + 1) * 4 * N].copy()), ptr(exp))
        got = srv.tap_galois(CID, level, ct)
        assert np.array_equal(got, exp), f"level {level}"


def test_expand_matches_oracle(cfg1):
    o, srv = cfg1["o"], cfg1["srv"]
    q = o.query(CLIENT_SEED, 1, cfg1["sk"], 4242)
    exp = o.expand(q, cfg1["gk"])
    got = srv.tap_expand(CID, q)
    assert np.array_equal(got, exp)


def test_convert_matches_oracle(cfg1):
    o, srv = cfg1["o"], cfg1["srv"]
    q = o.query(CLIENT_SEED, 2, cfg1["sk"], 1234)
    leaves = o.expand(q, cfg1["gk"])
    exp = o.convert(leaves, cfg1["rkn"])
    got = srv.tap_convert(CID, leaves)
    assert np.array_equal(got, exp)


def test_first_dim_matches_oracle(cfg1):
    o, srv = cfg1["o"], cfg1["srv"]
    rng = np.random.default_rng(3)
    sel = rng.integers(0, o.q, size=(o.fst, 2 * N), dtype=np.uint64)
    exp = o.first_dim(cfg1["db"], sel)
    got = srv.tap_first_dim(sel)
    assert np.array_equal(got, exp)


def test_mux_matches_oracle(cfg1):
    o, srv = cfg1["o"], cfg1["srv"]
    rng = np.random.default_rng(4)
    cts = rng.integers(0, o.q, size=(32, 2 * N), dtype=np.uint64)
    for bit in (0, 1):
        g = o.rgsw(CLIENT_SEED, 77 + bit, cfg1["sk"], bit)
        exp = o.mux_level(cts, g)
        got = srv.tap_mux(cts, g)
        assert np.array_equal(got, exp), f"bit {bit}"


def test_mod_switch_matches_oracle(cfg1):
    o, srv = cfg1["o"], cfg1["srv"]
    rng = np.random.default_rng(5)
    x = rng.integers(0, o.q, size=2 * N, dtype=np.uint64)
    x[:8] = [0, 1, o.q - 1, o.q // 2, o.q // 2 + 1, 2**33, 2**33 + 1, 4294967296 * 3]
    exp = np.array([o.L.orc_mod_switch_one(int(v), o.q, o.small_q) for v in x], dtype=np.uint64)
    got = srv.tap_mod_switch(x)
    assert np.array_equal(got, exp)


def test_db_generate_matches_oracle(cfg1):
    o, srv = cfg1["o"], cfg1["srv"]
    db = cfg1["db"].reshape(N, o.num_pt)
    for pt in (0, 1, 255, 4097, o.num_pt - 1):
        assert np.array_equal(srv.db_read_ntt(pt), db[:, pt]), f"pt {pt}"


@pytest.mark.parametrize("idx", [0, 4242, 8191])
def test_answer_bit_exact_and_decrypts(cfg1, idx):
    o, srv = cfg1["o"], cfg1["srv"]
    q = o.query(CLIENT_SEED, 100 + idx, cfg1["sk"], idx)
    exp, _ = o.answer(cfg1["db"], cfg1["gk"], cfg1["rkn"], q)
    got = srv.answer(CID, q)
    assert np.array_equal(got, exp)
    pt = o.decrypt_small(cfg1["sk"], got)
    assert np.array_equal(pt, o.db_plaintext(DB_SEED, 0, idx))
    packed = srv.pack_response(got)
    assert packed == o.pack(exp)
    t = srv.stage_times()
    assert t["total"] > 0


def test_unknown_client_raises(cfg1):
    from onionpirv2_amd import OnionKeyError
    q = np.zeros(2 * N, np.uint64)
    with pytest.raises(OnionKeyError):
        cfg1["srv"].answer(999, q)


def test_small_shape_generic_kernel_and_upload():
    """fst = 12 (not a multiple of 16) takes the generic first-dim kernel;
    the database arrives through the plaintext upload path.  Split and fused
    first multiplexer levels."""
    from onionpirv2_amd import GpuServer
    o = Oracle(fst=12, other=4)
    rng = np.random.default_rng(9)
    pts = rng.integers(0, o.t, size=(o.num_pt, N), dtype=np.uint64)
    pts[0, :4] = o.t - 1
    with GpuServer(12, 4) as srv:
        srv.db_upload_plain(0, pts)
        db = np.zeros(o.num_pt * N, np.uint64)
        for i in range(o.num_pt):
            db.reshape(N, o.num_pt)[:, i] = o.plain_to_ntt(pts[i])
        for i in (0, 5, o.num_pt - 1):
            assert np.array_equal(srv.db_read_ntt(i), db.reshape(N, o.num_pt)[:, i])
        sk = o.keygen(21)
        gk, rk = o.galois_keys(21, sk), o.rgsw_key(21, sk)
        srv.set_galois_key(1, gk)
        srv.set_rgsw_key(1, rk)
        for idx in (0, 7, o.num_pt - 1):
            q = o.query(21, idx, sk, idx)
            exp, _ = o.answer(db, gk, o.rgsw_key_ntt(rk), q)
            got = srv.answer(1, q)
            assert np.array_equal(got, exp)
            assert np.array_equal(o.decrypt_small(sk, got), pts[idx])
        # fuse_min = 1: the generic kernel's NTT-form outputs go straight into
        # a fused first multiplexer level
        srv.set_tuning(fuse_min=1)
        assert np.array_equal(srv.answer(1, q), exp)


@pytest.mark.parametrize("fst,other", [(1024, 2), (64, 1), (12, 4), (33, 8)])
def test_edge_shapes_single_and_batch(fst, other):
    """Shapes at the edges of the path, single queries and a batch of 5,
    against the oracle: fst = 1024 (expansion height 11 = log2 N, the
    tallest tree), other = 1 (one dimension: no conversion, no multiplexer),
    fst = 12 and 33 (first dimensions off the matrix cores' multiple of 32:
    the batch takes the 64-bit VALU tile).  At fst = 1024 the noise of an
    11-level expansion summed over 1024 products exceeds the decryption
    bound at both plaintext moduli (the oracle's responses do not decrypt
    either), so there only the bits are compared."""
    from onionpirv2_amd import GpuServer
    o = Oracle(fst=fst, other=other)
    db = o.db_build(31, 0)
    sk = o.keygen(77)
    gk, rk = o.galois_keys(77, sk), o.rgsw_key(77, sk)
    rkn = o.rgsw_key_ntt(rk)
    rng = np.random.default_rng(fst * 100 + other)
    idx = [0, o.num_pt - 1] + [int(i) for i in rng.integers(0, o.num_pt, size=3)]
    qs = np.stack([o.query(77, 600 + k, sk, i) for k, i in enumerate(idx)])
    exp = [o.answer(db, gk, rkn, q)[0] for q in qs]
    with GpuServer(fst, other) as srv:
        srv.db_generate(31, 0)
        srv.set_galois_key(3, gk)
        srv.set_rgsw_key(3, rk)
        for k in (0, 1):
            got = srv.answer(3, qs[k])
            assert np.array_equal(got, exp[k]), f"single query {k}"
            if fst <= 512:
                assert np.array_equal(o.decrypt_small(sk, got), o.db_plaintext(31, 0, idx[k]))
        got = srv.answer_batch([3] * len(idx), qs)
        for k in range(len(idx)):
            assert np.array_equal(got[k], exp[k]), f"batch query {k}"


def test_first_dim_wraps_mod_2_128():
    """All database words and selection values q-1: 512 * (q-1)^2 >= 2^128,
    so the reference's uint128 accumulator wraps; the device must agree."""
    from onionpirv2_amd import GpuServer
    o = Oracle(fst=512, other=2)
    big = np.full(N, o.q - 1, np.uint64)
    coeff = o.intt(big)  # selection poly whose NTT is all q-1
    sel = np.tile(np.concatenate([coeff, coeff]), (o.fst, 1))
    # database plaintexts whose NTT words are all q-1 are not reachable from
    # coefficients < t, so compare the raw first-dimension arithmetic on a
    # database of lifted t-1 plaintexts plus the closed form for all-(q-1).
    with GpuServer(512, 2) as srv:
        pts = np.full((o.num_pt, N), o.t - 1, np.uint64)
        srv.db_upload_plain(0, pts)
        db = np.zeros((N, o.num_pt), np.uint64)
        lifted = o.plain_to_ntt(np.full(N, o.t - 1, np.uint64))
        db[:, :] = lifted[:, None]
        wraps = sum(1 for v in lifted if 512 * int(v) * (o.q - 1) >= 2**128)
        assert wraps > 100  # the uint128 sum overflows in many slots
        exp = o.first_dim(db.reshape(-1), sel)
        got = srv.tap_first_dim(sel)
        assert np.array_equal(got, exp)
        # the batched kernels (matrix-core digit products, then the 64-bit tile)
        rng = np.random.default_rng(8)
        sel2 = rng.integers(0, o.q, size=(o.fst, 2 * N), dtype=np.uint64)
        batch = np.stack([sel, sel2, sel])
        for mfma_min, planes, var in ((2, 1, 0), (2, 1, 2), (2, 0, 0), (0, 0, 0)):
            srv.set_tuning(fd_mfma_min=mfma_min, db_planes=planes, fd_mfma_variant=var)
            gb = srv.tap_first_dim_batch(batch)
            assert np.array_equal(gb[0], exp) and np.array_equal(gb[2], exp), f"fd_mfma_min {mfma_min} var {var}"
            assert np.array_equal(gb[1], o.first_dim(db.reshape(-1), sel2))


def test_partial_upload_updates_the_plane_copy():
    """Plaintexts 21 .. 50 rewritten after db_generate: the records they share
    with untouched plaintexts (16 plaintexts each) are rebuilt, so the
    matrix-core product from the plane copy still matches the oracle."""
    from onionpirv2_amd import GpuServer
    o = Oracle(fst=64, other=4)
    rng = np.random.default_rng(31)
    lo, n = 21, 30
    pts = rng.integers(0, o.t, size=(n, N), dtype=np.uint64)
    db = o.db_build(3, 0).reshape(N, o.num_pt).copy()
    for i in range(n):
        p = pts[i]
        db[:, lo + i] = o.plain_to_ntt(p)
    sel = rng.integers(0, o.q, size=(3, o.fst, 2 * N), dtype=np.uint64)
    with GpuServer(64, 4) as srv:
        assert srv.get_tuning()["db_planes"] == 1
        srv.db_generate(3, 0)
        srv.db_upload_plain(lo, pts)
        srv.set_tuning(fd_mfma_min=2)
        got = srv.tap_first_dim_batch(sel)
        for b in range(3):
            assert np.array_equal(got[b], o.first_dim(db.reshape(-1), sel[b])), f"query {b}"
        srv.set_tuning(fd_mfma_variant=2)
        assert np.array_equal(srv.tap_first_dim_batch(sel), got)
        srv.set_tuning(db_planes=0)
        assert np.array_equal(srv.tap_first_dim_batch(sel), got)


@pytest.mark.parametrize("fst,other,nq", [(1024, 8, 3), (128, 64, 40), (96, 16, 20)])
def test_persistent_first_dim_matches_oracle(fst, other, nq):
    """The persistent matrix-core product (fd_mfma_variant = 2) walks several
    tiles per workgroup and reduces each tile's outputs during the next tile's
    steps.  fst = 1024: 32 steps per tile, more than its 16 outputs;
    fst = 128: 4 steps, the other 12 outputs reduced at the tile's end.
    nq = 3 takes one column block of 32 (122 of its 128 columns masked) and
    four row blocks (other = 8: 120 of 128 rows masked); nq = 40 takes four
    column blocks and one row block.  fst = 96 (an odd number of steps) is
    outside the persistent kernel's shapes, so the per-tile kernel runs.
    It matches the oracle and the per-tile kernel."""
    from onionpirv2_amd import GpuServer
    o = Oracle(fst=fst, other=other)
    db = o.db_build(6, 0)
    rng = np.random.default_rng(fst + nq)
    sel = rng.integers(0, o.q, size=(nq, fst, 2 * N), dtype=np.uint64)
    sel[0, :, :4] = o.q - 1
    with GpuServer(fst, other) as srv:
        srv.db_generate(6, 0)
        srv.set_tuning(fd_mfma_min=2, fd_mfma_variant=2)
        got = srv.tap_first_dim_batch(sel)
        for b in sorted({0, nq - 1}):
            assert np.array_equal(got[b], o.first_dim(db, sel[b])), f"query {b}"
        srv.set_tuning(fd_mfma_variant=0)
        assert np.array_equal(srv.tap_first_dim_batch(sel), got)


@pytest.mark.parametrize("shape", [(64, 2), (64, 4), (256, 16)])
def test_first_mux_level_from_ntt_form(shape):
    """A batch whose first multiplexer level runs from the matrix-core
    product's NTT-form outputs (fuse_min = 1) answers as the oracle does, and
    as the path that transforms every first-dimension output first
    (fuse_min = 2^30).  half = 1, 2 and 8 take both workgroup-to-XCD mappings;
    at other = 2 the first level is the only one.
    The levels after it keep NTT-form ciphertexts while they are wide
    (k_mux_fused_ntt) and return to coefficient form for the narrow split
    kernels: fuse_min = 5 / 10 / 20 put that switch after the first, second
    or third level.  The batch of 3 takes the first level one query per
    workgroup, a batch of 4 two queries per workgroup, and a batch of 8 at
    half = 8 two queries per workgroup in groups of 8 queries per XCD."""
    from onionpirv2_amd import GpuServer
    fst, other = shape
    o = Oracle(fst=fst, other=other)
    db = o.db_build(5, 0)
    keys = []
    for seed in (41, 42, 43):
        sk = o.keygen(seed)
        keys.append((seed, sk, o.galois_keys(seed, sk), o.rgsw_key(seed, sk)))
    with GpuServer(fst, other) as srv:
        srv.db_generate(5, 0)
        for cid, (seed, sk, gk, rk) in enumerate(keys):
            srv.set_galois_key(cid, gk)
            srv.set_rgsw_key(cid, rk)
        idx = [3, o.num_pt - 1, (o.num_pt * 2) // 3]
        qs = np.stack([o.query(keys[i][0], 100 + i, keys[i][1], idx[i]) for i in range(3)])
        srv.set_tuning(fd_mfma_min=2, fuse_min=1)
        fused = srv.answer_batch([0, 1, 2], qs)
        srv.set_tuning(fuse_min=1 << 30)
        assert np.array_equal(srv.answer_batch([0, 1, 2], qs), fused)
        for fm in (5, 10, 20):
            srv.set_tuning(fuse_min=fm)
            assert np.array_equal(srv.answer_batch([0, 1, 2], qs), fused), f"fuse_min {fm}"
        for i in (0, 2):
            seed, sk, gk, rk = keys[i]
            exp, _ = o.answer(db, gk, o.rgsw_key_ntt(rk), qs[i])
            assert np.array_equal(fused[i], exp), f"query {i}"
            assert np.array_equal(o.decrypt_small(sk, fused[i]), o.db_plaintext(5, 0, idx[i]))
        # an even batch takes the first level two queries per workgroup
        q4 = o.query(keys[0][0], 104, keys[0][1], 7)
        exp4, _ = o.answer(db, keys[0][2], o.rgsw_key_ntt(keys[0][3]), q4)
        srv.set_tuning(fuse_min=1)
        even = srv.answer_batch([0, 1, 2, 0], np.concatenate([qs, q4[None]]))
        assert np.array_equal(even[:3], fused)
        assert np.array_equal(even[3], exp4)
        c8 = [i % 3 for i in range(8)]
        q8 = np.stack([o.query(keys[c][0], 200 + i, keys[c][1], (5 * i + 1) % o.num_pt) for i, c in enumerate(c8)])
        got8 = srv.answer_batch(c8, q8)
        for i, c in enumerate(c8):
            assert np.array_equal(got8[i], srv.answer(c, q8[i])), f"batch of 8, query {i}"
        exp8, _ = o.answer(db, keys[1][2], o.rgsw_key_ntt(keys[1][3]), q8[7])
        assert np.array_equal(got8[7], exp8)


@pytest.mark.parametrize("l,l_key", [(3, 4), (13, 11), (5, 16), (4, 10), (6, 10)])
def test_other_gadget_lengths(l, l_key):
    """Gadget lengths other than (5, 10) decide which folds apply: the
    multiplexer's (l - 1) B < 60 holds for l = 3 (B = 20), 4 (B = 15), 5 and
    6 (B = 10), not for l = 13 (B = 5: 60), which keeps coefficient-form
    levels; its NTT-form kernel transforms the l - 1 digits two at a time, so
    l = 4 and 6 also take the single transform after the pairs; the conversion
    key's holds for l_key = 4 (B = 15), not for 11 (B = 6) or 16 (B = 4, and
    32 products per 128-bit sum, the most the bound allows), which take the
    unfolded conversion of NTT-form leaves (top rows copied, then returned to
    coefficient form in place).  Fused (fuse_min = 1) and split shapes give the
    oracle's bits for the expansion, conversion, a single answer and a batch."""
    from onionpirv2_amd import GpuServer
    fst, other = 64, 4
    o = Oracle(fst=fst, other=other, l=l, l_key=l_key)
    db = o.db_build(9, 0)
    keys = []
    for seed in (61, 62):
        sk = o.keygen(seed)
        keys.append((seed, sk, o.galois_keys(seed, sk), o.rgsw_key(seed, sk)))
    idx = [o.num_pt - 1, 70]
    qs = np.stack([o.query(keys[i][0], 300 + i, keys[i][1], idx[i]) for i in range(2)])
    rkn = [o.rgsw_key_ntt(k[3]) for k in keys]
    exp = [o.answer(db, keys[i][2], rkn[i], qs[i])[0] for i in range(2)]
    leaves = o.expand(qs[0], keys[0][2])
    conv = o.convert(leaves, rkn[0])
    with GpuServer(fst, other, gsw_l=l, gsw_l_key=l_key) as srv:
        srv.db_generate(9, 0)
        for cid, (seed, sk, gk, rk) in enumerate(keys):
            srv.set_galois_key(cid, gk)
            srv.set_rgsw_key(cid, rk)
        for fm in (1, 1 << 30):
            srv.set_tuning(fuse_min=fm, fd_mfma_min=2)
            assert np.array_equal(srv.tap_expand(0, qs[0]), leaves), f"fuse_min {fm}"
            assert np.array_equal(srv.tap_convert(0, leaves), conv), f"fuse_min {fm}"
            assert np.array_equal(srv.answer(1, qs[1]), exp[1]), f"fuse_min {fm}"
            got = srv.answer_batch([0, 1], qs)
            for i in range(2):
                assert np.array_equal(got[i], exp[i]), f"fuse_min {fm}, query {i}"


@pytest.mark.parametrize("shape", [(32, 4), (64, 8), (256, 16), (1024, 2)])
def test_expansion_leaves_in_ntt_form(shape):
    """Fused expansion levels (fuse_min = 1: all of them) run in NTT form from
    the first on and write NTT-form leaves: the Galois substitution as a slot
    permutation and x^-k as a slot factor.
    Tree heights 6, 7, 9 and 11 (k = 32 .. 1024 at the last level).  The
    expansion tap (leaves returned to coefficient form), a single answer, and
    batches through the matrix-core and the VALU first dimension (NTT-form
    selection vectors, NTT-form conversion top rows) all give the oracle's
    bits, and the same bits as coefficient-form leaves (fuse_min = 2^30)."""
    from onionpirv2_amd import GpuServer
    fst, other = shape
    o = Oracle(fst=fst, other=other)
    db = o.db_build(8, 0)
    keys = []
    for seed in (51, 52):
        sk = o.keygen(seed)
        keys.append((seed, sk, o.galois_keys(seed, sk), o.rgsw_key(seed, sk)))
    idx = [o.num_pt - 1, o.num_pt // 3]
    qs = np.stack([o.query(keys[i][0], 200 + i, keys[i][1], idx[i]) for i in range(2)])
    exp = [o.answer(db, keys[i][2], o.rgsw_key_ntt(keys[i][3]), qs[i])[0] for i in range(2)]
    with GpuServer(fst, other) as srv:
        srv.db_generate(8, 0)
        for cid, (seed, sk, gk, rk) in enumerate(keys):
            srv.set_galois_key(cid, gk)
            srv.set_rgsw_key(cid, rk)
        srv.set_tuning(fuse_min=1)
        assert np.array_equal(srv.tap_expand(0, qs[0]), o.expand(qs[0], keys[0][2]))
        assert np.array_equal(srv.answer(1, qs[1]), exp[1])
        for mfma_min in (2, 0):
            srv.set_tuning(fd_mfma_min=mfma_min)
            got = srv.answer_batch([0, 1], qs)
            for i in range(2):
                assert np.array_equal(got[i], exp[i]), f"fd_mfma_min {mfma_min}, query {i}"
        srv.set_tuning(fuse_min=1 << 30, fd_mfma_min=2)
        assert np.array_equal(srv.answer_batch([0, 1], qs), got)


@pytest.mark.parametrize("nq", [2, 3, 16, 17, 33, 64])
def test_first_dim_batch_matches_oracle(cfg1, nq):
    """The batched first-dimension kernels (int8 matrix cores by default, the
    64-bit VALU tile with fd_mfma_min = 0) against the oracle, query by query.
    nq = 2..16 / 17..32 / 33..64 take 1 / 2 / 4 column blocks of 32; other =
    32 leaves 96 of the 128 rows of a workgroup masked."""
    o, srv = cfg1["o"], cfg1["srv"]
    rng = np.random.default_rng(100 + nq)
    sel = rng.integers(0, o.q, size=(nq, o.fst, 2 * N), dtype=np.uint64)
    sel[0, :, :8] = o.q - 1
    sel[-1, :3] = 0
    got = srv.tap_first_dim_batch(sel)
    check = sorted({0, nq // 2, nq - 1})
    for b in check:
        assert np.array_equal(got[b], o.first_dim(cfg1["db"], sel[b])), f"query {b}"
    saved = srv.get_tuning()
    try:
        srv.set_tuning(fd_mfma_min=0)
        assert np.array_equal(srv.tap_first_dim_batch(sel), got)
        for var, planes in ((0, 1), (1, 1), (2, 1), (0, 0), (1, 0)):
            srv.set_tuning(fd_mfma_min=2, fd_mfma_variant=var, db_planes=planes)
            assert np.array_equal(srv.tap_first_dim_batch(sel), got), f"variant {var} planes {planes}"
    finally:
        srv.set_tuning(**saved)


@pytest.mark.parametrize("nq", [3, 17, 40])
def test_answer_batch_matches_single_and_oracle(cfg1, nq):
    """A batch of nq queries from nq different clients (each with its own keys)
    gives, query by query, the single-query response; three of them are also
    checked against the oracle.  nq = 3 / 17 / 40 take the 32 / 64 / 128-column
    variants of the batched first-dimension kernel."""
    o, srv = cfg1["o"], cfg1["srv"]
    rng = np.random.default_rng(nq)
    idx = rng.choice(o.num_pt, size=nq, replace=False)
    ids, qs, keys = [], [], []
    for i in range(nq):
        cid = 1000 + i
        seed = 500 + i
        sk = o.keygen(seed)
        gk, rk = o.galois_keys(seed, sk), o.rgsw_key(seed, sk)
        srv.set_galois_key(cid, gk)
        srv.set_rgsw_key(cid, rk)
        ids.append(cid)
        qs.append(o.query(seed, i, sk, int(idx[i])))
        keys.append((sk, gk, rk))
    got = srv.answer_batch(ids, np.stack(qs))
    for i in range(nq):
        assert np.array_equal(got[i], srv.answer(ids[i], qs[i])), f"query {i}"
    for i in (0, nq // 2, nq - 1):
        sk, gk, rk = keys[i]
        exp, _ = o.answer(cfg1["db"], gk, o.rgsw_key_ntt(rk), qs[i])
        assert np.array_equal(got[i], exp)
        assert np.array_equal(o.decrypt_small(sk, got[i]), o.db_plaintext(DB_SEED, 0, int(idx[i])))
    for cid in ids:
        srv.drop_client(cid)


@pytest.mark.parametrize("tuning", [dict(fuse_min=1), dict(fuse_min=1 << 30),
                                    dict(fd_rows_b1=64), dict(fd_tile_rows=2),
                                    dict(fd_b1_variant=1), dict(fd_b1_variant=2),
                                    dict(fd_b1_variant=3), dict(fd_mfma_min=0),
                                    dict(fd_mfma_min=6), dict(fd_mfma_variant=0),
                                    dict(fd_mfma_variant=1), dict(fd_mfma_variant=2),
                                    dict(db_planes=0),
                                    dict(db_planes=0, fd_mfma_variant=1),
                                    dict(ntt_radix=4), dict(ntt_radix=4, fuse_min=1)])
def test_tuning_choices_are_bit_identical(cfg1, tuning):
    """Every launch-shape choice (fused vs split external products, first-dim
    tile shapes, transform pass schedules) gives the oracle's bits, for a
    single query and batches of 5 and 4 (the first multiplexer level takes
    one query per workgroup for an odd batch, two for an even one)."""
    o, srv = cfg1["o"], cfg1["srv"]
    saved = srv.get_tuning()
    try:
        srv.set_tuning(**tuning)
        rng = np.random.default_rng(77)
        cts = rng.integers(0, o.q, size=(32, 2 * N), dtype=np.uint64)
        g = o.rgsw(CLIENT_SEED, 91, cfg1["sk"], 1)
        assert np.array_equal(srv.tap_mux(cts, g), o.mux_level(cts, g))
        q = o.query(CLIENT_SEED, 300, cfg1["sk"], 777)
        leaves = o.expand(q, cfg1["gk"])
        assert np.array_equal(srv.tap_expand(CID, q), leaves)
        assert np.array_equal(srv.tap_convert(CID, leaves), o.convert(leaves, cfg1["rkn"]))
        exp, _ = o.answer(cfg1["db"], cfg1["gk"], cfg1["rkn"], q)
        assert np.array_equal(srv.answer(CID, q), exp)
        got = srv.answer_batch([CID] * 5, np.stack([q] * 5))
        for i in range(5):
            assert np.array_equal(got[i], exp)
        got = srv.answer_batch([CID] * 4, np.stack([q] * 4))
        for i in range(4):
            assert np.array_equal(got[i], exp)
    finally:
        srv.set_tuning(**saved)


def test_batch_split_above_one_pass(cfg1):
    """A call of 70 queries runs as 64 + 6 (batch_split = 0) or 35 + 35
    (batch_split = 1): the same responses, each the single-query answer; two
    also checked against the oracle.  The autotuner's batch-split axis runs for
    a size above 64 and leaves results unchanged."""
    o, srv = cfg1["o"], cfg1["srv"]
    saved = srv.get_tuning()
    rng = np.random.default_rng(70)
    idx = rng.choice(o.num_pt, size=70, replace=False)
    qs = np.stack([o.query(CLIENT_SEED, 2000 + i, cfg1["sk"], int(idx[i])) for i in range(70)])
    ids = [CID] * 70
    try:
        srv.set_tuning(batch_split=0)
        a = srv.answer_batch(ids, qs)
        srv.set_tuning(batch_split=1)
        b = srv.answer_batch(ids, qs)
        assert np.array_equal(a, b)
        for i in (0, 34, 35, 63, 64, 69):
            assert np.array_equal(a[i], srv.answer(CID, qs[i])), f"query {i}"
        for i in (5, 66):
            exp, _ = o.answer(cfg1["db"], cfg1["gk"], cfg1["rkn"], qs[i])
            assert np.array_equal(a[i], exp)
        chosen = srv.autotune([1, 70], reps=1)
        assert chosen["batch_split"] in (0, 1) and chosen["ntt_radix"] in (8, 4)
        assert np.array_equal(srv.answer_batch(ids, qs), a)
    finally:
        srv.set_tuning(**saved)


def test_autotune_keeps_results_bit_identical(cfg1):
    """The autotuner only changes launch shapes: after it, single and batched
    answers still equal the oracle's."""
    o, srv = cfg1["o"], cfg1["srv"]
    saved = srv.get_tuning()
    try:
        chosen = srv.autotune([1, 5], reps=2)
        assert chosen == srv.get_tuning()
        q = o.query(CLIENT_SEED, 501, cfg1["sk"], 99)
        exp, _ = o.answer(cfg1["db"], cfg1["gk"], cfg1["rkn"], q)
        assert np.array_equal(srv.answer(CID, q), exp)
        got = srv.answer_batch([CID] * 5, np.stack([q] * 5))
        assert all(np.array_equal(g, exp) for g in got)
    finally:
        srv.set_tuning(**saved)


def test_measurement_taps(cfg1):
    """The profiler-calibration copy moves what it reports, and the VALU
    issue-rate probe runs and refuses a grid larger than its sink."""
    from onionpirv2_amd._capi import OnionError
    srv = cfg1["srv"]
    assert srv.tap_copy(1 << 20) == 1 << 20
    ms = srv.tap_valu_probe(256, 64)
    assert 0 < ms < 1e3
    with pytest.raises(OnionError):
        srv.tap_valu_probe(1 << 30, 1)
