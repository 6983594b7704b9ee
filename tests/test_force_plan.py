"""The force kernel's column planner, on the CPU (lulesh_force_plan_of: the
planner the engine runs, without a GPU).

Whatever plan it picks (tile shape, even z-runs or the paired plan, any CU
count), every node plane of every node column of every subdomain must be
advanced by exactly one run: a plane covered twice or never would leave the
march's results wrong at that size only.  The tiles must also fit the
kernel (at most 256 element slots, a node patch of at most 320 slots per
plane), carry the first-tile flag exactly on each subdomain's origin tile,
and a paired plan must fit one launch round (at most two workgroups per
CU).  The plans measured in the design notes are pinned as they are.
"""
import numpy as np
import pytest

FIRST_TILE = 2


@pytest.fixture(scope="module")
def lib():
    """the library loaded without a GPU (host-only entry points)"""
    import lulesh_amd
    lulesh_amd.load()
    return lulesh_amd


def check_plan(lib, nx, tp, cus):
    kc, paired, t = lib.force_plan_of(nx, tp, cus)
    s1 = nx + 1
    ndom = tp ** 3
    assert kc >= 1
    dom, i0, j0, k0, ex, ey, flags, kn = t.T
    assert np.all((dom >= 0) & (dom < ndom))
    assert np.all(kn >= 1) and np.all(k0 >= 0) and np.all(k0 + kn <= s1)
    assert np.all(ex * ey <= 256) and np.all((ex + 1) * (ey + 1) <= 320)
    assert np.all((ex >= 2) & (ey >= 2))
    # one tile shape per subdomain (all subdomains of a group are alike)
    assert len(set(zip(ex.tolist(), ey.tolist()))) == 1
    e, f = int(ex[0]), int(ey[0])
    # columns start on the (ex-1, ey-1) node grid
    assert np.all(i0 % (e - 1) == 0) and np.all(j0 % (f - 1) == 0)
    assert np.all(i0 < s1) and np.all(j0 < s1)
    origin = (i0 == 0) & (j0 == 0) & (k0 == 0)
    assert np.array_equal(flags == FIRST_TILE, origin)
    assert np.all((flags == 0) | (flags == FIRST_TILE))
    assert origin.sum() == ndom
    # each node column's runs partition [0, s1) once
    cover = np.zeros((ndom, s1, s1, s1), dtype=np.int32)
    for d, a, b, k, m in zip(dom, i0, j0, k0, kn):
        cover[d, k:k + m, b:min(s1, b + f - 1), a:min(s1, a + e - 1)] += 1
    assert np.all(cover == 1), (nx, tp, cus, int(cover.min()), int(cover.max()))
    if paired:
        assert len(t) <= 2 * cus
    return kc, paired, t


@pytest.mark.parametrize("cus", [256, 304, 80, 8, 1])
def test_every_plane_once_single_domain(lib, cus):
    for nx in list(range(1, 121)) + [128, 150, 175, 200, 256]:
        check_plan(lib, nx, 1, cus)


@pytest.mark.parametrize("tp", [2, 3, 4])
def test_every_plane_once_decomposed(lib, tp):
    for cus in (256, 80):
        for nx in (1, 2, 3, 4, 7, 10, 16, 23, 31, 45, 48):
            check_plan(lib, nx, tp, cus)


def test_plans_of_the_design_notes(lib):
    # -s 100 on 256 CUs: the paired plan, 510 workgroups (48 columns of
    # 18 x 14); 2x2x2 of -s 45: 512; -s 64: 407; -s 200: the even split
    kc, paired, t = lib.force_plan_of(100, 1, 256)
    assert paired and len(t) == 510 and (t[0, 4], t[0, 5]) == (18, 14)
    kc, paired, t = lib.force_plan_of(45, 2, 256)
    assert paired and len(t) == 512
    kc, paired, t = lib.force_plan_of(64, 1, 256)
    assert paired and len(t) == 407
    kc, paired, t = lib.force_plan_of(200, 1, 256)
    assert not paired


def test_bad_arguments(lib):
    for args in ((0, 1, 256), (8, 0, 256), (8, 5, 256), (8, 1, 0), (2000, 1, 256)):
        with pytest.raises(lib.LuleshError):
            lib.force_plan_of(*args)
