"""svs_verify_noalloc (device r and v_enc, fixed scratch, no allocation)
against svs_verify (verify_response, scorer-bntm/src/crypto.rs:1013-1046)
through the C ABI: the same verdicts, honest and tampered, on two calls in a
row with the second on a smaller list, so that scratch left by the larger
call must not leak into the smaller one.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
P = O.P


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _field(rng, size):
    return rng.integers(0, P, size=size, dtype=np.uint64)


def _random_cluster(rng, m):
    rowptr = np.zeros(m + 1, np.uint32)
    return dict(m=m, m_enc=_field(rng, m * 1024), a_l=_field(rng, m * 128),
                h=_field(rng, m * 128), m_plain=None, s_rowptr=rowptr,
                s_idx=np.zeros(0, np.uint32), s_sign=np.zeros(0, np.int8))


def _both(ctx, idx, lst, r, v_enc, seed):
    import svs
    r_dev = svs.DeviceArray.from_host(ctx, np.ascontiguousarray(r, np.uint64))
    v_dev = svs.DeviceArray.from_host(ctx, np.ascontiguousarray(v_enc, np.uint64))
    try:
        got = idx.verify_noalloc(lst, r_dev.ptr(), v_dev.ptr(), seed)
    finally:
        r_dev.close()
        v_dev.close()
    want = idx.verify(lst, r, v_enc, seed)
    return got, want


def test_verify_noalloc_matches_verify():
    import svs
    ctx = svs.Context(0)
    l_sub = O.generate_subspace(bytes([1] * 32))
    rng = np.random.default_rng(0x7A11)
    sizes = [20000, 300]  # the second call runs on the smaller list
    idx = svs.GpuIndex(ctx, len(sizes), l_sub)
    cls = []
    for i, m in enumerate(sizes):
        cl = _random_cluster(rng, m)
        idx.upload_list(i, cl)
        cls.append(cl)
    for rep in range(2):
        for i, m in enumerate(sizes):
            v_enc = _field(rng, 1024)
            r = idx.compute_products(i, v_enc)
            seed = bytes([(7 * i + rep) & 0xFF] * 32)
            say("noalloc honest: list", i, "rows", m, "pass", rep)
            got, want = _both(ctx, idx, i, r, v_enc, seed)
            assert got == want == -1, (got, want)
            t = r.copy()
            t[m // 2] = (int(t[m // 2]) + 1) % P
            say("noalloc tampered: list", i, "rows", m, "pass", rep)
            got, want = _both(ctx, idx, i, t, v_enc, seed)
            assert got == want == 0, (got, want)
    # an error orthogonal to u_0 is caught at trial 1, on the small list
    m = sizes[1]
    v_enc = _field(rng, 1024)
    r = O.matvec(cls[1]["m_enc"], m, 1024, v_enc)
    seed = bytes([35] * 32)
    u0 = O.u_draws(seed, m)
    ea = 5
    eb = (-ea * u0[3] * pow(u0[11], P - 2, P)) % P
    r[3] = (int(r[3]) + ea) % P
    r[11] = (int(r[11]) + eb) % P
    say("noalloc trial-1 tamper")
    got, want = _both(ctx, idx, 1, r, v_enc, seed)
    assert got == want == 1 == O.verify_response(r, cls[1]["m_enc"], m, v_enc, seed)
    # a non-canonical v_enc is an argument error, as in svs_verify
    bad = v_enc.copy()
    bad[5] = np.uint64(P)
    from svs import _lib as L
    b_dev = svs.DeviceArray.from_host(ctx, bad)
    r_dev = svs.DeviceArray.from_host(ctx, r)
    with pytest.raises(L.GpuError) as e:
        idx.verify_noalloc(1, r_dev.ptr(), b_dev.ptr(), seed)
    assert e.value.code == L.SVS_ERR_ARG
    b_dev.close()
    r_dev.close()
    idx.close()
