"""Batched zone operations on the device, bit-exact against the oracle's restatement.

Inputs follow the zone-kernel microbenchmark: seeded canonical zones (point-cloud hulls, loosened,
closed), d in {4, 8, 16, 32}, L/U ~ U[0, 1000], inclusion pairs half built as B inside A.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import binding
from ptctcl_amd import hip
from ptctcl_amd import zonegen as zg

pytestmark = pytest.mark.gpu
DIMS = [4, 8, 16, 32]
N = 3000


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def run(op, d, a, b=None, arg=None, L=None, U=None):
    n = a.shape[0]
    A = dev(a)
    out = torch.zeros_like(A)
    flag = torch.zeros(n, dtype=torch.int32, device="cuda:0")
    hip.zone_op(op, d, A, None if b is None else dev(b), None if arg is None else dev(arg),
                None if L is None else dev(L), None if U is None else dev(U), out, flag)
    torch.cuda.synchronize()
    return out.cpu().numpy(), flag.cpu().numpy()


@pytest.mark.parametrize("d", DIMS)
def test_close(d):
    rng = np.random.default_rng(100 + d)
    z = zg.unclose(zg.hull_zones(N, d, rng), d, rng)
    eo, ef = binding.zone_op(hip.ZOP_CLOSE, d, z)
    go, gf = run(hip.ZOP_CLOSE, d, z)
    assert np.array_equal(ef, gf)
    assert np.array_equal(eo, go)


@pytest.mark.parametrize("d", DIMS)
@pytest.mark.parametrize("op", [hip.ZOP_UP, hip.ZOP_RESET])
def test_up_reset(d, op):
    rng = np.random.default_rng(200 + d)
    z = zg.hull_zones(N, d, rng)
    arg = rng.integers(1, d, size=N).astype(np.int32) if op == hip.ZOP_RESET else None
    eo, _ = binding.zone_op(op, d, z, arg=arg)
    go, _ = run(op, d, z, arg=arg)
    assert np.array_equal(eo, go)


@pytest.mark.parametrize("d", DIMS)
@pytest.mark.parametrize("op", [hip.ZOP_DOWN, hip.ZOP_FREE])
def test_down_free(d, op):
    rng = np.random.default_rng(250 + d)
    z = zg.hull_zones(N, d, rng)
    arg = None
    if op == hip.ZOP_FREE:
        arg = (rng.integers(1, 1 << min(d - 1, 30), size=N).astype(np.int64) << 1).astype(np.int32)
    eo, _ = binding.zone_op(op, d, z, arg=arg)
    go, _ = run(op, d, z, arg=arg)
    assert np.array_equal(eo, go)


@pytest.mark.parametrize("d", DIMS)
def test_constrain(d):
    rng = np.random.default_rng(300 + d)
    z = zg.hull_zones(N, d, rng)
    c = zg.constraints(z, d, rng)
    eo, ef = binding.zone_op(hip.ZOP_CONSTRAIN, d, z, arg=c)
    go, gf = run(hip.ZOP_CONSTRAIN, d, z, arg=c)
    assert np.array_equal(ef, gf)
    ne = ef == 1     # an empty result's matrix is unspecified
    assert 0 < ne.sum() < N
    assert np.array_equal(eo[ne], go[ne])


@pytest.mark.parametrize("d", DIMS)
def test_extra_lu_plus(d):
    rng = np.random.default_rng(400 + d)
    z = zg.hull_zones(N, d, rng)
    L, U = zg.bounds(N, d, rng)
    eo, _ = binding.zone_op(hip.ZOP_EXTRA_LU, d, z, L=L, U=U)
    go, _ = run(hip.ZOP_EXTRA_LU, d, z, L=L, U=U)
    assert np.array_equal(eo, go)


@pytest.mark.parametrize("d", DIMS)
def test_includes_and_alu(d):
    rng = np.random.default_rng(500 + d)
    z = zg.hull_zones(N, d, rng)
    a, b = zg.pairs_included(z, d, rng)
    L, U = zg.bounds(N, d, rng, neg_inf_frac=0.1)
    _, ef = binding.zone_op(hip.ZOP_INCLUDES, d, a, b)
    _, gf = run(hip.ZOP_INCLUDES, d, a, b)
    assert np.array_equal(ef, gf) and 0 < ef.sum() < N
    for x, y in ((a, b), (b, a)):
        _, ef = binding.zone_op(hip.ZOP_ALU_LE, d, x, y, L=L, U=U)
        _, gf = run(hip.ZOP_ALU_LE, d, x, y, L=L, U=U)
        assert np.array_equal(ef, gf)


@pytest.mark.parametrize("d", DIMS)
def test_try_union(d):
    rng = np.random.default_rng(600 + d)
    z = zg.hull_zones(N, d, rng)
    a, b = zg.pairs_union(z, d, rng)
    eo, ef = binding.zone_op(hip.ZOP_UNION, d, a, b)
    go, gf = run(hip.ZOP_UNION, d, a, b)
    assert np.array_equal(ef, gf) and 0 < ef.sum() < N
    assert np.array_equal(eo, go)


@pytest.mark.parametrize("d", DIMS)
def test_full_size_properties(d):
    """2^20 zones at every d (a pool of distinct seeded zones tiled on the device, as in the
    microbenchmark): size-independent properties checked on the device (close idempotent on
    canonical zones, Extra_LU+ with bounds above every constant the identity, a_LU and inclusion
    reflexive, union with itself convex and the zone itself)."""
    n = 1 << 20
    pool = {4: 1 << 14, 8: 1 << 14, 16: 1 << 12, 32: 1 << 10}[d]
    rng = np.random.default_rng(700 + d)
    Z = dev(zg.hull_zones(pool, d, rng)).repeat(n // pool, 1)
    out = torch.empty_like(Z)
    flag = torch.zeros(n, dtype=torch.int32, device="cuda:0")
    hip.zone_op(hip.ZOP_CLOSE, d, Z, out=out, flag=flag)
    assert int(flag.min()) == 1 and torch.equal(out, Z)
    big = torch.full((n, d), 1 << 20, dtype=torch.int32, device="cuda:0")
    hip.zone_op(hip.ZOP_EXTRA_LU, d, Z, L=big, U=big, out=out, flag=flag)
    assert torch.equal(out, Z)
    Lp, Up = zg.bounds(pool, d, rng)
    L, U = dev(Lp).repeat(n // pool, 1), dev(Up).repeat(n // pool, 1)
    flag.zero_()
    hip.zone_op(hip.ZOP_ALU_LE, d, Z, Z, L=L, U=U, flag=flag)
    assert int(flag.min()) == 1
    flag.zero_()
    hip.zone_op(hip.ZOP_UNION, d, Z, Z, out=out, flag=flag)
    assert int(flag.min()) == 1 and torch.equal(out, Z)
    torch.cuda.synchronize()


@pytest.mark.parametrize("d", DIMS)
@pytest.mark.parametrize("shape", ["1,1", "2,3", "4,1000000"])
def test_register_kernels_launch_shapes(d, shape, monkeypatch):
    """Close and Extra_LU+ at every launch shape the autotuner may pick: one wavefront walking all
    zone groups (loading the next group while closing one), a few workgroups, one group per
    wavefront; bit-exact against the oracle."""
    monkeypatch.setenv("HORA_HIP_ZOP_SHAPE", shape)
    rng = np.random.default_rng(800 + d)
    z = zg.unclose(zg.hull_zones(301, d, rng), d, rng)
    eo, ef = binding.zone_op(hip.ZOP_CLOSE, d, z)
    go, gf = run(hip.ZOP_CLOSE, d, z)
    assert np.array_equal(ef, gf) and np.array_equal(eo, go)
    z = zg.hull_zones(301, d, rng)
    L, U = zg.bounds(301, d, rng)
    eo, _ = binding.zone_op(hip.ZOP_EXTRA_LU, d, z, L=L, U=U)
    go, _ = run(hip.ZOP_EXTRA_LU, d, z, L=L, U=U)
    assert np.array_equal(eo, go)


@pytest.mark.parametrize("d", [4, 16])
def test_autotuned_and_graph_captured(d):
    """The first call with 2^17 zones tunes the register kernels' launch shape; the same calls
    captured in a HIP graph (through torch's graph capture of the current stream) and replayed
    three times, with eager work and synchronisation between replays, give the eager answer each
    time, bit-exact against the oracle on a sample."""
    n = 1 << 17
    pool = 1 << 12
    rng = np.random.default_rng(900 + d)
    zp = zg.unclose(zg.hull_zones(pool, d, rng), d, rng)
    Z = dev(zp).repeat(n // pool, 1)
    ref = torch.empty_like(Z)
    flag_ref = torch.zeros(n, dtype=torch.int32, device="cuda:0")
    hip.zone_op(hip.ZOP_CLOSE, d, Z, out=ref, flag=flag_ref)      # tunes
    torch.cuda.synchronize()
    eo, ef = binding.zone_op(hip.ZOP_CLOSE, d, zp)
    assert np.array_equal(ref[:pool].cpu().numpy(), eo) and np.array_equal(flag_ref[:pool].cpu().numpy(), ef)
    out = torch.zeros_like(Z)
    flag = torch.zeros(n, dtype=torch.int32, device="cuda:0")
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        hip.zone_op(hip.ZOP_CLOSE, d, Z, out=out, flag=flag, stream=s.cuda_stream)   # warm-up off-graph
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        hip.zone_op(hip.ZOP_CLOSE, d, Z, out=out, flag=flag, stream=torch.cuda.current_stream().cuda_stream)
    for rep in range(3):
        out.fill_(7)
        flag.fill_(-1)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, ref) and torch.equal(flag, flag_ref), rep
@pytest.mark.parametrize("case", __import__("extra_cases").EXTRA_CASES, ids=lambda c: c[0])
def test_extra_lu_plus_known_answers(case):
    """Hand-derived Extra_LU+ answers (tests/extra_cases.py), one per clause, on the device."""
    _, d, z, L, U, want = case
    go, _ = run(hip.ZOP_EXTRA_LU, d, np.asarray([z], np.int32), L=np.asarray([L], np.int32),
                U=np.asarray([U], np.int32))
    assert go[0].tolist() == want
