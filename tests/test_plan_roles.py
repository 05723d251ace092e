"""Role layout of the persistent create launches, checked on the host.

The fused pass (p2g_dev_create_pass) and the batched pass run encode
workgroups (producers) beside hash workgroups; the packet hashes wait for the
encode's column bands.  Every encode workgroup must come before every
workgroup that waits on it, so that a grid which is only partly resident
(another process holding CUs) dispatches its producers first, and the grid
never exceeds what the device holds at once.  The reference has no such
launch: its workers own disjoint outputs and never wait on each other
(src/creator.rs:965-995), which is the property this layout keeps."""
import pytest

import par2rust_amd as p2

SHAPES = [
    # (slices, recovery, slice size): the headline set, small and ragged sets,
    # the many-files set, one 32 GiB set, a one-slice set
    (1027, 200, 524288),
    (1024, 10, 4096),
    (37, 1, 4),
    (4674, 1000, 65536),
    (32768, 3000, 1 << 20),
    (1, 1, 64),
    (2000, 65535, 512),
]


def _check_layout(roles, grid_cap):
    assert 0 < len(roles) <= grid_cap
    enc = [b for b, r in enumerate(roles) if r == p2.ROLE_ENCODE]
    sl = [b for b, r in enumerate(roles) if r == p2.ROLE_SLICE]
    pk = [b for b, r in enumerate(roles) if r == p2.ROLE_PACKET]
    assert enc, "no encode workgroup"
    assert enc == list(range(len(enc))), "encode workgroups are not the lowest indices"
    # slice hashes (wait for nothing) next, packet hashes (wait for bands) last
    assert sl == list(range(len(enc), len(enc) + len(sl)))
    assert pk == list(range(len(enc) + len(sl), len(roles)))
    return len(enc), len(sl), len(pk)


@pytest.mark.parametrize("n,R,S", SHAPES)
def test_batched_pass_roles_producers_first(n, R, S):
    roles = p2.plan_roles("batch", 256, 256, n, R, S)
    e, h2, h3 = _check_layout(roles, 256)
    assert h3 >= 1 and h2 >= 1


@pytest.mark.parametrize("n,R,S", [s for s in SHAPES if s[0] <= 2048])
def test_fused_pass_roles_producers_first(n, R, S):
    roles = p2.plan_roles("pass", 256, 256, n, R, S)
    e, h2, h3 = _check_layout(roles, 256)
    assert h3 == min((R + 63) // 64, 8) and h2 == min((n + 63) // 64, 32)


def test_headline_pass_layout():
    # 1027 slices, 200 blocks: 17 slice-hash and 4 packet-hash workgroups
    roles = p2.plan_roles("pass", 256, 256, 1027, 200, 524288)
    assert _check_layout(roles, 256) == (235, 17, 4)


@pytest.mark.parametrize("kind", ["pass", "batch"])
def test_grid_never_exceeds_resident_workgroups(kind):
    # a device that holds fewer workgroups than it has CUs: the grid shrinks
    for resident in (200, 97):
        roles = p2.plan_roles(kind, 256, resident, 1027, 200, 524288)
        _check_layout(roles, resident)
    if kind == "batch":
        _check_layout(p2.plan_roles(kind, 256, 64, 1027, 200, 524288), 64)
    else:
        # 21 hash workgroups need a grid of more than 4 x 21: the fused pass
        # is not planned and the set runs as separate launches instead
        with pytest.raises(p2.Par2Error):
            p2.plan_roles(kind, 256, 64, 1027, 200, 524288)


def test_fused_pass_refuses_sets_it_cannot_hold():
    with pytest.raises(p2.Par2Error):
        p2.plan_roles("pass", 256, 256, 4096, 200, 65536)  # more than 2048 slices
    with pytest.raises(p2.Par2Error):
        p2.plan_roles("pass", 16, 16, 1027, 200, 524288)  # too few CUs for the hash roles
