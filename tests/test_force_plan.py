"""The force kernel's column planner, on the CPU (lulesh_force_plan_of: the
planner the engine runs, without a GPU).

Whatever plan it picks (tile shape, even z-runs or the paired plan, the
z-chunk a batch asks for, any CU count), every node plane of every node
column of every subdomain must be advanced by exactly one run: a plane
covered twice or never would leave the march's results wrong at that size
only.  The tiles must also fit the kernel (at most 256 element slots, a
node patch of at most 320 slots per plane), carry the first-tile flag
exactly on each subdomain's origin tile, and a paired plan must fit one
launch round (at most two workgroups per CU).  The plans measured in the
design notes are pinned as they are.
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


def check_plan(lib, nx, tp, cus, kchunk=0):
    """nx, tp: one group, or one entry per group of an engine"""
    nxs = list(np.atleast_1d(nx))
    tps = list(np.broadcast_to(np.asarray(tp), (len(nxs),)))
    kc, paired, t = lib.force_plan_of(nxs, tps, cus, kchunk)
    assert kc >= 1 and (kchunk == 0 or kc == kchunk)
    assert not (paired and (len(nxs) > 1 or kchunk))
    # subdomain d -> its mesh size, numbered over the groups in order
    dom_nx = np.concatenate([[n] * (p ** 3) for n, p in zip(nxs, tps)])
    dom, i0, j0, k0, ex, ey, flags, kn = t.T
    assert np.all((dom >= 0) & (dom < dom_nx.size))
    s1 = dom_nx[dom] + 1
    assert np.all(kn >= 1) and np.all(k0 >= 0) and np.all(k0 + kn <= s1)
    assert np.all(ex * ey <= 256) and np.all((ex + 1) * (ey + 1) <= 320)
    assert np.all((ex >= 2) & (ey >= 2))
    assert np.all(i0 % (ex - 1) == 0) and np.all(j0 % (ey - 1) == 0)
    assert np.all(i0 < s1) and np.all(j0 < s1)
    origin = (i0 == 0) & (j0 == 0) & (k0 == 0)
    assert np.array_equal(flags == FIRST_TILE, origin)
    assert np.all((flags == 0) | (flags == FIRST_TILE))
    assert origin.sum() == dom_nx.size
    # one tile shape per subdomain; each node column's runs partition [0, s1)
    for d in range(dom_nx.size):
        sel = dom == d
        assert len(set(zip(ex[sel].tolist(), ey[sel].tolist()))) == 1, d
        n1 = int(dom_nx[d]) + 1
        cover = np.zeros((n1, n1, n1), dtype=np.int32)
        for a, b, k, m, e, f in zip(i0[sel], j0[sel], k0[sel], kn[sel], ex[sel], ey[sel]):
            cover[k:k + m, b:min(n1, b + f - 1), a:min(n1, a + e - 1)] += 1
        assert np.all(cover == 1), (nxs, tps, cus, kchunk, d, int(cover.min()), int(cover.max()))
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


@pytest.mark.parametrize("kchunk", [0, 1, 3, 4, 8, 32])
def test_every_plane_once_ensembles(lib, kchunk):
    # engines of several groups (batches, servers): seeded mixes like the
    # benchmark's ensembles, with a decomposed member among them
    rng = np.random.default_rng(11 + kchunk)
    for trial in range(4):
        n = int(rng.integers(2, 40))
        nxs = rng.integers(1, 65, n)
        tps = np.ones(n, dtype=int)
        tps[0] = 2
        for cus in (256, 80):
            check_plan(lib, nxs, tps, cus, kchunk)


def test_one_group_with_a_chunk_given(lib):
    # a z-chunk asked for (the batch autotuner's candidates) takes the even
    # split with that chunk, never the paired plan
    for nx in (30, 64, 100):
        for kchunk in (4, 8):
            kc, paired, _ = check_plan(lib, nx, 1, 256, kchunk)
            assert kc == kchunk and not paired


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
    with pytest.raises(lib.LuleshError):
        lib.force_plan_of(8, 1, 256, -1)
