"""GPU parity of EMVP's score GEMV and Tiptoe's SimplePIR answer against the
oracle (bit-exact). Mirrors scorer-emvp/src/crypto.rs:463-507 and
scorer-tiptoe/src/pir/simplepir.rs:293-376."""
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
P = O.P


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import svs
    return svs.Context(0)


def test_emvp_partials_match_oracle(ctx, golden):
    from svs.schemes import EmvpCluster
    e = golden["emvp"]
    rng = np.random.default_rng(e["seed"])
    m_hat = rng.integers(0, P, size=(e["m"], 1292), dtype=np.uint64)
    q_hat = rng.integers(0, P, size=1292, dtype=np.uint64)
    say("emvp golden")
    cl = EmvpCluster(m_hat, ctx)
    assert [str(int(x)) for x in cl.compute_products(q_hat)] == e["out"]
    for m in (1, 7, 63, 64, 65, 257):
        say("emvp m =", m)
        mh = rng.integers(0, P, size=(m, 1292), dtype=np.uint64)
        mh[0, :] = P - 1
        cl = EmvpCluster(mh, ctx)
        qs = rng.integers(0, P, size=(5, 1292), dtype=np.uint64)
        qs[0] = P - 1
        got = cl.compute_products_batch(qs)
        want = O.emvp_compute_products_batch(mh, m, qs)
        assert np.array_equal(got, want)
    assert EmvpCluster(np.zeros((0, 1292), np.uint64), ctx).compute_products(q_hat) is None


def test_emvp_dev_grid_stride_matches_oracle(ctx):
    """Device-pointer entry at a size whose work units outnumber the grid
    (every block takes several row tiles x column groups), a ragged last
    row tile, and the host-buffer entry on the same inputs."""
    import svs
    from svs.schemes import EmvpCluster
    rng = np.random.default_rng(12)
    m, nq = 5001, 16
    mh = rng.integers(0, P, size=(m, 1292), dtype=np.uint64)
    mh[m - 1, :] = P - 1
    qs = rng.integers(0, P, size=(nq, 1292), dtype=np.uint64)
    qs[nq - 1] = P - 1
    cl = EmvpCluster(mh, ctx)
    want = O.emvp_compute_products_batch(mh, m, qs)
    say("emvp dev m =", m, "nq =", nq)
    qd = svs.DeviceArray.from_host(ctx, qs)
    out = svs.DeviceArray(ctx, (nq, 76 * m), np.uint64)
    assert cl.compute_products_dev(nq, qd.ptr(), out.ptr())
    assert np.array_equal(out.to_host(), want)
    assert np.array_equal(cl.compute_products_batch(qs[:3]), want[:3])
    qd.close()
    out.close()
    cl.close()


def test_tiptoe_kat(ctx, golden):
    from svs.schemes import TiptoeDatabase
    t = golden["tiptoe_kat"]
    db = TiptoeDatabase(np.array(t["m"], np.uint64), ctx)
    x = np.array(t["x"], np.uint64)
    for col in range(2):
        assert db.server_answer(np.ascontiguousarray(x[:, col])).tolist() == \
            [row[col] for row in t["mx"]]


def test_tiptoe_u64_and_u32_match_oracle(ctx, golden):
    from svs.schemes import TiptoeDatabase
    t = golden["tiptoe"]
    rng = np.random.default_rng(t["seed"])
    mdb = rng.integers(0, 32, size=(t["rows"], t["cols"]), dtype=np.uint64)
    ct = rng.integers(0, 1 << 64, size=t["cols"], dtype=np.uint64)
    say("tiptoe golden u64")
    assert [str(int(x)) for x in TiptoeDatabase(mdb, ctx).server_answer(ct)] == t["ans64"]
    db32 = TiptoeDatabase(mdb.astype(np.uint32), ctx)
    assert [str(int(x)) for x in db32.server_answer((ct & np.uint64(0xFFFFFFFF)).astype(np.uint32))] \
        == t["ans32"]


@pytest.mark.parametrize("rows,cols,nq", [(16, 64, 1), (100, 192, 3), (257, 1024, 17),
                                          (64, 8192, 64)])
def test_tiptoe_mod32_kernels(ctx, rows, cols, nq):
    from svs.schemes import TiptoeDatabase
    from svs._lib import KERNEL_MFMA, KERNEL_VALU
    rng = np.random.default_rng(rows * 7 + nq)
    mdb = rng.integers(0, 32, size=(rows, cols), dtype=np.uint32)
    mdb[0] = 31
    ct = rng.integers(0, 1 << 32, size=(nq, cols), dtype=np.uint32)
    ct[0] = 0xFFFFFFFF
    db = TiptoeDatabase(mdb, ctx)
    assert db.stored() == (rows * cols, 1)  # 5-bit entries stored one byte each
    want = np.stack([O.tiptoe_answer_u32(mdb, rows, cols, c) for c in ct])
    for kernel in (KERNEL_VALU, KERNEL_MFMA, -1):
        say("tiptoe", rows, cols, nq, "kernel", kernel)
        assert np.array_equal(db.server_answer(ct, kernel), want)


def test_tiptoe_wide_entries_use_u32_store(ctx):
    from svs.schemes import TiptoeDatabase
    rng = np.random.default_rng(3)
    mdb = rng.integers(0, 1 << 32, size=(40, 100), dtype=np.uint32)
    ct = rng.integers(0, 1 << 32, size=(2, 100), dtype=np.uint32)
    db = TiptoeDatabase(mdb, ctx)
    assert db.stored() == (40 * 100 * 4, 4)
    want = np.stack([O.tiptoe_answer_u32(mdb, 40, 100, c) for c in ct])
    assert np.array_equal(db.server_answer(ct), want)
