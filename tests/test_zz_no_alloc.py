"""The batched min-plus product that allocates nothing
(mi355_minplus_gemm_f32_batched_noalloc), through the C ABI, against the
batched entry point it varies (mi355_minplus_gemm_f32_batched) and the
oracle: bit-exact on two calls in a row, the second with a smaller batch, and
the capacity limits rejected with a clear error before any launch."""
import ctypes as C

import numpy as np
import pytest

import mi355sssp as M
from oracle import orc

pytestmark = pytest.mark.gpu


def _batch(rng, count, lo, hi):
    As, Bs = [], []
    for _ in range(count):
        m, n, k = (int(x) for x in rng.integers(lo, hi, 3))
        As.append((rng.integers(1, 1 << 24, (m, k)) * 2.0 ** -24).astype(np.float32))
        Bs.append((rng.integers(1, 1 << 24, (k, n)) * 2.0 ** -24).astype(np.float32))
    return As, Bs


def _run(fn, As, Bs):
    bufs = [(M.DeviceBuffer.from_array(a), M.DeviceBuffer.from_array(b), M.DeviceBuffer(a.shape[0] * b.shape[1] * 4))
            for a, b in zip(As, Bs)]
    cnt = len(bufs)
    arr = lambda xs: (C.c_void_p * cnt)(*xs)  # noqa: E731
    ints = lambda xs: (C.c_int * cnt)(*xs)  # noqa: E731
    Ms = [a.shape[0] for a in As]
    Ns = [b.shape[1] for b in Bs]
    Ks = [a.shape[1] for a in As]
    rc = fn(cnt, arr([x[0].ptr.value for x in bufs]), arr([x[1].ptr.value for x in bufs]),
            arr([x[2].ptr.value for x in bufs]), ints(Ms), ints(Ns), ints(Ks), ints(Ks), ints(Ns), ints(Ns), None)
    M._abi.check(rc)
    M._abi.check(M.engine().mi355_stream_sync(None))
    return [x[2].to_array(np.float32, Ms[i] * Ns[i]).reshape(Ms[i], Ns[i]) for i, x in enumerate(bufs)]


def test_noalloc_batch_matches_the_allocating_batch_twice():
    eng = M.engine()
    rng = np.random.default_rng(77)
    for count, lo, hi in ((48, 8, 300), (7, 3, 90)):  # the second call is smaller
        print(f"no-alloc batch: {count} products, sizes in [{lo}, {hi})", flush=True)
        As, Bs = _batch(rng, count, lo, hi)
        want = _run(eng.mi355_minplus_gemm_f32_batched, As, Bs)
        got = _run(eng.mi355_minplus_gemm_f32_batched_noalloc, As, Bs)
        for i in range(count):
            assert got[i].tobytes() == want[i].tobytes(), f"product {i} differs"
        for i in (0, count - 1):
            m, k = As[i].shape
            n = Bs[i].shape[1]
            ref = np.asarray(orc.minplus_gemm(As[i], Bs[i], m, n, k), np.float32).reshape(m, n)
            assert got[i].tobytes() == ref.tobytes()


def test_noalloc_batch_rejects_batches_beyond_its_array():
    eng = M.engine()
    dummy = M.DeviceBuffer(64)

    def call(count, m, n, k):
        ptrs = (C.c_void_p * count)(*([dummy.ptr.value] * count))
        ints = lambda x: (C.c_int * count)(*([x] * count))  # noqa: E731
        return eng.mi355_minplus_gemm_f32_batched_noalloc(count, ptrs, ptrs, ptrs, ints(m), ints(n), ints(k),
                                                          ints(k), ints(n), ints(n), None)

    # more than 4096 products, and one product with more tiles than the plan area
    for args in ((4097, 4, 4, 4), (1, 100000, 100000, 100000)):
        rc = call(*args)
        assert rc == -6, (args, rc)  # MI355_ECAPACITY
        assert b"no-allocation" in eng.mi355_last_error()
    assert eng.mi355_minplus_gemm_f32_batched_noalloc(0, None, None, None, None, None, None, None, None, None,
                                                      None) == 0
