"""Host planning of the packed launch (CPU): strip shapes, replica blocks,
task splitting and the co-residency bound, on the finite-size-scaling set
of Ising L in {16, ..., 256} and XY L in {16, ..., 256} with 48 temperatures."""

import pytest

from core.packed import choose_shape, group_shapes, plan_pack, split_tasks

ISING_L = (16, 24, 32, 48, 64, 96, 128, 256)
XY_L = (16, 32, 64, 128, 256)


def _lds(code, L, H, K):
    # the kernel's layout: RNG rows 1..H+1 plus state rows (mc_strip.h); PT area ignored here
    per = (H + 1) * L * 16 + ((H + 3) * L if code == 0 else (H + 3) * L * 8 + (H + 1) * L * 4)
    return K * ((per + 15) // 16 * 16) + 4096


def test_shapes_respect_the_kernel_constraints():
    for threads in (256, 512):
        for L in ISING_L:
            for H, K in group_shapes(L, 48, threads):
                assert L % H == 0 and H % 2 == 0 and H >= 4
                assert 48 % K == 0 and K & (K - 1) == 0 and K <= threads // 16
                assert (threads // (L // 2)) // K >= 1 and K * L <= threads
    assert group_shapes(1024, 48, 512) == []  # L > threads


def test_chosen_tasks_are_near_the_target():
    for L in ISING_L:
        H, K = choose_shape(L, 48, 256, 2048)
        assert 0.5 <= K * H * L / 2048 <= 2.25, (L, H, K)
    assert choose_shape(256, 48, 256, 4096) == (16, 1)
    assert choose_shape(16, 48, 256, 4096) == (16, 16)


@pytest.mark.parametrize("n,T", [(10, 3), (7, 7), (1, 4), (100, 9)])
def test_split_covers_every_task_once_evenly(n, T):
    parts = split_tasks(n, T)
    assert len(parts) == -(-n // T)
    assert [f for f, _ in parts] == [sum(c for _, c in parts[:i]) for i in range(len(parts))]
    assert sum(c for _, c in parts) == n
    counts = [c for _, c in parts]
    assert max(counts) - min(counts) <= 1 and max(counts) <= T


@pytest.mark.parametrize("capacity", [256, 1024, 4096])
def test_finite_size_set_plan_fits_the_capacity(capacity):
    groups = [(0, L, 48, (lambda H, K, L=L: _lds(0, L, H, K))) for L in ISING_L]
    groups += [(2, L, 48, (lambda H, K, L=L: _lds(2, L, H, K))) for L in XY_L]
    plan = plan_pack(groups, 256, 2048, lambda lds: capacity)
    work = plan["work"]
    assert len(work) <= capacity
    T = plan["tasks_per_workgroup"]
    assert all(1 <= c <= T for _, _, c in work)
    # smallest T that fits: one fewer task per workgroup would not
    ntasks = [(48 // K) * (L // H) for (_, L, _, _), (H, K) in zip(groups, plan["shapes"])]
    if T > 1:
        assert sum(-(-n // (T - 1)) for n in ntasks) > capacity
    for g, n in enumerate(ntasks):
        rows = [(f, c) for gg, f, c in work if gg == g]
        assert sum(c for _, c in rows) == n and rows[0][0] == 0
    # a group's workgroups are consecutive
    seen = [g for g, _, _ in work]
    assert seen == sorted(seen)


def test_plan_rejects_what_cannot_run():
    groups = [(0, 16, 48, lambda H, K: 10 ** 9)]
    with pytest.raises(ValueError):
        plan_pack(groups, 256, 2048, lambda lds: 100)
    with pytest.raises(ValueError):
        choose_shape(18, 48, 256, 2048, lds_of=lambda H, K: 10 ** 9)
