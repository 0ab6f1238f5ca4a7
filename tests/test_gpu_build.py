"""GPU parity of the one-time build (growTree ... reduceBounds, sortKeys,
buildBVH, initCamera/updateCamera) through the C-ABI against the oracle.

Bar: Keys, Order and the node topology (counts, primitive ids) bit-exact
with the reference-harness restatement (glibc libm); Nodes bit-exact when
built from identical P. The GPU evaluates the build's sin/cos/pow correctly
rounded (in double, rounded once); glibc's sinf/cosf are not correctly
rounded on ~1.4% of inputs, so B and P differ from the glibc restatement by
at most 1e-5 m and are bit-exact against the restatement with correctly
rounded sin/cos/pow (oracle generate_cr).
"""
import numpy as np
import pytest

from gpu_util import require_gpu, to_dev, to_host, ulp_diff
from oracle_lib import CAPACITY, OBJECT_COUNT, PS

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def px():
    require_gpu()
    import pyarbor

    return pyarbor


def _report(name, a, b):
    d = ulp_diff(a, b)
    print("%s: %d/%d words differ, max ulp %d" % (name, int((d > 0).sum()), d.size, int(d.max())))
    return d


@pytest.mark.parametrize("seed", [1, 7, 1788])
def test_generate_matches_oracle(px, oracle, seed):
    print("case: arbor_generate seed", seed, flush=True)
    dev = px.DeviceScene().generate(seed)
    ref = oracle.generate(seed)
    B, P = dev.host("B"), dev.host("P")
    _report("B vs glibc restatement", B, ref.B)
    _report("P vs glibc restatement", P, ref.P)
    assert np.abs(B - ref.B).max() <= 1e-5 and np.abs(P - ref.P).max() <= 1e-5
    assert np.array_equal(dev.host("Keys"), ref.Keys)
    assert np.array_equal(dev.host("Order"), ref.Order)
    N, R = dev.host("Nodes").reshape(-1, 8), ref.Nodes.reshape(-1, 8)
    assert np.array_equal(N[:, 3], R[:, 3]) and np.array_equal(N[:, 7], R[:, 7])  # topology, counts, ids
    assert N[1, 3] == OBJECT_COUNT
    cr = oracle.generate_cr(seed)
    for name, got, want in (("B", B, cr.B), ("P", P, cr.P), ("Nodes", N.reshape(-1), cr.Nodes)):
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), name + " vs correctly rounded libm"
    C = dev.host("C")
    assert np.array_equal(C.view(np.uint32), oracle.init_camera(seed).view(np.uint32))


def test_build_from_identical_P_is_bit_exact(px, oracle):
    print("case: arbor_buildBVH on oracle P (seed 7)", flush=True)
    ref = oracle.generate(7)
    P = to_dev(ref.P)
    import torch

    K = torch.zeros(CAPACITY, dtype=torch.int32, device="cuda")
    O = torch.zeros_like(K)
    N = torch.zeros(CAPACITY * 2 * 8, dtype=torch.float32, device="cuda")
    px.call("arbor_buildBVH", P, K, O, N)
    assert np.array_equal(to_host(K, np.uint32), ref.Keys)
    assert np.array_equal(to_host(O, np.uint32), ref.Order)
    assert np.array_equal(to_host(N).view(np.uint32), ref.Nodes.view(np.uint32))


def test_reference_stage_sequence(px, oracle):
    """The 12-entry build sequence of engine.js:29-32 with 120 sortPass stages."""
    print("case: stage-by-stage build seed 1788", flush=True)
    import torch

    seed = 1788
    ref = oracle.generate(seed)
    cr = oracle.generate_cr(seed)
    s = px.DeviceScene()
    for level in range(5):
        px.call("arbor_growTree", s.B, level, seed)
    B = s.host("B")
    assert np.array_equal(B.view(np.uint32), cr.B.view(np.uint32))
    # makePrimitives from the (correctly rounded) oracle's B isolates the stage
    Bd = to_dev(cr.B)
    Pd = torch.zeros(CAPACITY * PS, dtype=torch.float32, device="cuda")
    px.call("arbor_makePrimitives", Bd, Pd, seed)
    assert np.array_equal(to_host(Pd).view(np.uint32), cr.P.view(np.uint32))
    # keys -> bitonic -> bounds -> reduce from the oracle's P
    P = to_dev(ref.P)
    px.call("arbor_makeKeys", P, s.Keys, s.Order)
    k0, o0 = oracle.make_keys(ref.P)
    assert np.array_equal(s.host("Keys"), k0) and np.array_equal(s.host("Order"), o0)
    span = 2
    while span <= CAPACITY:
        d = span // 2
        while d > 0:
            px.call("arbor_sortPass", s.Keys, s.Order, d, span)
            d //= 2
        span *= 2
    assert np.array_equal(s.host("Keys"), ref.Keys) and np.array_equal(s.host("Order"), ref.Order)
    px.call("arbor_makeBounds", P, s.Order, s.Nodes)
    first = CAPACITY // 2
    while first >= 1:
        px.call("arbor_reduceBounds", s.Nodes, first)
        first //= 2
    assert np.array_equal(s.host("Nodes").view(np.uint32), ref.Nodes.view(np.uint32))


def test_sortKeys_equals_sorted_pairs(px, oracle):
    print("case: arbor_sortKeys random inputs", flush=True)
    rng = np.random.default_rng(11)
    cases = [
        (rng.integers(0, 64, CAPACITY).astype(np.uint32), np.arange(CAPACITY, dtype=np.uint32)),
        (rng.integers(0, 2**32, CAPACITY, dtype=np.uint64).astype(np.uint32), rng.permutation(CAPACITY).astype(np.uint32)),
        (rng.integers(0, 16, CAPACITY).astype(np.uint32), rng.integers(0, 2**32, CAPACITY, dtype=np.uint64).astype(np.uint32)),
        (np.full(CAPACITY, 0xFFFFFFFF, np.uint32), np.arange(CAPACITY, dtype=np.uint32)[::-1].copy()),
    ]
    for keys, order in cases:
        k, o = keys.copy(), order.copy()
        oracle.sort_pairs(k, o)
        K, O = to_dev(keys), to_dev(order)
        px.call("arbor_sortKeys", K, O)
        assert np.array_equal(to_host(K, np.uint32), k)
        assert np.array_equal(to_host(O, np.uint32), o)


def test_camera_state_machine(px, oracle, golden):
    print("case: initCamera/updateCamera scripted inputs", flush=True)
    import torch

    g, crops = golden
    C = torch.zeros(64, dtype=torch.float32, device="cuda")
    px.call("arbor_initCamera", C, None, 7)
    assert np.array_equal(to_host(C), np.array(g["camera_init_7"], np.float32))
    for step, (I, want) in enumerate(zip(crops["camera_inputs"], crops["camera_states"])):
        px.call("arbor_updateCamera", C, to_dev(I), None, 1.0 / 60.0, 320, 240)
        got = to_host(C)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (step, np.nonzero(got != want))
