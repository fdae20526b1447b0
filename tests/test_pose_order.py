"""The summation order jr_pose's normal kernel reproduces, pinned on the CPU.

merge_objects normalises each object's transformed normals by the 2-norm of
the whole array, whose float32 sum of squares numpy adds (np.sum) in 8192-value
chunks taken in turn, each chunk summed pairwise: at most 128 values in eight
interleaved partial sums, longer ranges halved at n // 2 rounded down to a
multiple of 8.  k_pose_normals walks the same trees with explicit stacks (the
leaves left to right, then a post-order combine); `kernel_order` below
restates that walk step for step.  Both must equal np.sum bit for bit."""
import numpy as np

f32 = np.float32
CHUNK, LEAF = 8192, 128


def leaf_sum(a, s, n):
    if n < 8:
        r = f32(0.0)
        for i in range(n):
            r = f32(r + a[s + i])
        return r
    r = [a[s + j] for j in range(8)]
    i = 8
    while i < n - n % 8:
        for j in range(8):
            r[j] = f32(r[j] + a[s + i + j])
        i += 8
    res = f32(f32(f32(r[0] + r[1]) + f32(r[2] + r[3])) + f32(f32(r[4] + r[5]) + f32(r[6] + r[7])))
    for i in range(n - n % 8, n):
        res = f32(res + a[s + i])
    return res


def split(n):
    n2 = n // 2
    return n2 - n2 % 8


def kernel_order(a):
    """Returns (sum, number of leaves, deepest stack)."""
    n = len(a)
    leaves, deepest = [], 0
    for cs in range(0, n, CHUNK):
        stack = [(cs, min(CHUNK, n - cs))]
        while stack:
            s, ln = stack.pop()
            if ln <= LEAF:
                leaves.append((s, ln))
            else:
                n2 = split(ln)
                stack.append((s + n2, ln - n2))
                stack.append((s, n2))
            deepest = max(deepest, len(stack))
    sums = [leaf_sum(a, s, ln) for s, ln in leaves]
    total, nxt = f32(0.0), 0
    for cs in range(0, n, CHUNK):
        right, left, has = [], [], []
        ln = min(CHUNK, n - cs)
        while True:
            while ln > LEAF:
                n2 = split(ln)
                right.append(ln - n2)
                left.append(f32(0.0))
                has.append(False)
                ln = n2
            v = sums[nxt]
            nxt += 1
            while has and has[-1]:
                v = f32(left[-1] + v)
                right.pop(), left.pop(), has.pop()
            if not has:
                break
            left[-1], has[-1] = v, True
            ln = right[-1]
        total = f32(total + v)
    assert nxt == len(sums)
    return total, len(leaves), deepest


def test_kernel_order_equals_numpy_float32_sum():
    rng = np.random.default_rng(0)
    sizes = list(range(0, 300)) + [8192, 8193, 16384, 16385, 24577] + \
        [int(x) for x in rng.integers(300, 40000, 12)]
    for n in sizes:
        v = (rng.standard_normal(n) * 10 ** rng.uniform(-2, 2)).astype(f32)
        sq = (v * v).astype(f32)
        got, leaves, deepest = kernel_order(sq)
        want = np.sum(sq, dtype=f32) if n else f32(0.0)
        assert got.view(np.uint32) == np.asarray(want, f32).view(np.uint32), n
        # the kernel's shared-memory bounds: leaves <= n // 64 + chunks, stacks < 32
        assert leaves <= n // 64 + -(-n // CHUNK)
        assert deepest < 32


def test_a_plain_left_to_right_sum_would_differ():
    """The order matters: a sequential float32 sum of the same squares
    differs from np.sum on most of these arrays, so the test above pins it."""
    rng = np.random.default_rng(1)
    differ = 0
    for _ in range(20):
        sq = (rng.standard_normal(3000).astype(f32) ** 2).astype(f32)
        seq = f32(0.0)
        for x in sq:
            seq = f32(seq + x)
        differ += seq != np.sum(sq, dtype=f32)
    assert differ >= 10
