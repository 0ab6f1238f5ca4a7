"""Structure of the SAH trace hierarchy (the device layout that
arbor-cuda_amd/csrc/accel.hip writes), checked against the heap BVH it was built
from: every primitive of the heap appears once, with its heap leaf box; every
child box of an interior node is the exact union of that child's own
children's boxes; the depth in the header is the deepest leaf's and within
the walk's 32-entry stack; the slot table maps each primitive to its heap
leaf (the order exact ties resolve in)."""
import numpy as np

CAPACITY = 32768


def check_accel_structure(words, nodes, order):
    """words: u32 words of the hierarchy; nodes: heap Nodes (f32, NODE_COUNT
    x 8); order: Order (slot -> primitive id). Returns the depth."""
    words = np.ascontiguousarray(words, np.uint32)
    f = words.view(np.float32)
    i = words.view(np.int32)
    n_nodes, n_prims, depth = int(f[8]), int(f[9]), int(f[7])
    nodes = np.asarray(nodes, np.float32).reshape(-1, 8)
    order = np.asarray(order).view(np.int32)
    leaves = nodes[CAPACITY:]
    used = leaves[:, 3] > 0
    assert n_prims == int(used.sum()) and n_nodes == n_prims - 1
    assert 1 <= depth <= 32
    heap_box = {int(order[s]): leaves[s] for s in np.nonzero(used)[0]}
    seen_prims, seen_nodes = set(), set()

    def child(q, k):  # box (lo, hi) and ref of child k of interior node q
        base = 16 + 16 * q + 8 * k
        return f[base:base + 3], f[base + 4:base + 7], int(i[base + 3])

    stack = [(int(i[3]), f[0:3], f[4:7], 0)]
    max_level = 0
    while stack:
        ref, lo, hi, level = stack.pop()
        max_level = max(max_level, level)
        if ref < 0:
            pid = ~ref
            assert pid not in seen_prims and pid in heap_box
            seen_prims.add(pid)
            hb = heap_box[pid]
            assert np.array_equal(lo, hb[0:3]) and np.array_equal(hi, hb[4:7])
            continue
        assert 0 <= ref < n_nodes and ref not in seen_nodes
        seen_nodes.add(ref)
        l_lo, l_hi, l_ref = child(ref, 0)
        r_lo, r_hi, r_ref = child(ref, 1)
        assert np.array_equal(lo, np.minimum(l_lo, r_lo)) and np.array_equal(hi, np.maximum(l_hi, r_hi))
        stack.append((l_ref, l_lo, l_hi, level + 1))
        stack.append((r_ref, r_lo, r_hi, level + 1))
    assert len(seen_prims) == n_prims and len(seen_nodes) == n_nodes
    assert max_level == depth
    slots = words[16 + 16 * (CAPACITY - 1):][:CAPACITY]
    for s in np.nonzero(used)[0]:
        assert slots[order[s]] == s
    return depth


class LeafSet:
    """Nodes (heap layout, leaf records from CAPACITY on) and Order of a
    synthetic leaf set, on the host."""


def skewed_leaf_set(shape):
    """Leaf sets built to defeat the surface-area heuristic: every box the
    same ("identical": no split separates them), centres spaced
    exponentially along one axis ("exponential": each SAH split would peel
    off one leaf), nested boxes growing geometrically ("nested"), and all
    CAPACITY slots used ("full", random boxes). Ids are a seeded permutation
    of the used slots."""
    cap = CAPACITY
    rng = np.random.default_rng(5)
    n = {"identical": 20000, "exponential": 30000, "nested": 4000, "full": cap}[shape]
    nodes = np.zeros((2 * cap, 8), np.float32)
    leaves = nodes[cap:]
    if shape == "identical":
        c = np.zeros((n, 3), np.float32)
    elif shape == "exponential":
        c = np.zeros((n, 3), np.float32)
        c[:, 0] = np.float32(1.0009) ** np.arange(n, dtype=np.float32)
    else:
        c = rng.uniform(-10, 10, (n, 3)).astype(np.float32)
    half = np.float32(0.01)
    if shape == "nested":  # each box holds every smaller one: the heuristic would peel one per level
        g = np.float32(1.004) ** np.arange(n, dtype=np.float32)
        c = np.zeros((n, 3), np.float32)
        c[:, 0] = g
        half = (g * np.float32(2.0))[:, None]
    leaves[:n, 0:3] = c - half
    leaves[:n, 4:7] = c + half
    leaves[:n, 3] = 1.0
    order = np.arange(cap, dtype=np.uint32)
    perm = rng.permutation(n).astype(np.uint32)
    order[:n] = perm
    leaves[:n, 7] = perm.astype(np.float32)
    s = LeafSet()
    s.Nodes = np.ascontiguousarray(nodes.reshape(-1))
    s.Order = order
    return s
