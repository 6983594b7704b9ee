"""Single-domain parity of the HIP path against the CPU oracle.

The oracle in LO_CBRT_EXACT mode uses the correctly rounded cube root the
kernels compute; everything else is the reference's arithmetic in the
reference's order, so the GPU state must match it bit for bit.  Against the
oracle in libm mode (= the compiled reference) the bar is the origin-energy
tolerance of 1e-10 relative.
"""
import numpy as np
import pytest

from conftest import assert_bitwise

pytestmark = pytest.mark.gpu

NODE = ["x", "y", "z", "xd", "yd", "zd"]
ELEM = ["e", "p", "q", "v", "ss", "volo", "delv", "vdov", "arealg", "vnew",
        "delx_xi", "delx_eta", "delx_zeta"]


def test_probe_cbrt_matches_exact(amd, oracle_mod):
    rng = np.random.default_rng(1)
    x = np.concatenate([
        rng.uniform(1e-9, 1e-2, 20000), rng.uniform(0.5, 8.0, 20000),
        np.exp(rng.uniform(-700, 700, 20000)), -rng.uniform(1e-3, 1e3, 2000),
        np.array([1.0, 8.0, 27.0, 0.125, 2.0, 7.999999999999999, 1e-300, 5e-324, 1.7e308]),
    ])
    got = amd.probe_cbrt(x)
    want = np.array([oracle_mod.cbrt_exact(v) for v in x])
    assert_bitwise(got, want, "cbrt")


def test_probe_cbrt_hard_cases(amd, oracle_mod):
    # inputs whose cube roots lie next to a rounding midpoint (the nearest
    # doubles to ((2Y+1)/2^53)^3 and their neighbours, at many exponents)
    # and exact cubes: the cases the fast rounding decision must hand to
    # the exact midpoint test or get right by itself
    from fractions import Fraction
    rng = np.random.default_rng(3)
    xs = []
    for _ in range(4000):
        y = int(rng.integers(1 << 52, 1 << 53))
        mid = float(Fraction((2 * y + 1) ** 3, 2 ** 159))
        sh = 3 * int(rng.integers(-300, 300))
        for v in (mid, np.nextafter(mid, 0.0), np.nextafter(mid, 16.0)):
            xs.append(np.ldexp(v, sh))
    for _ in range(4000):
        a = int(rng.integers(1, 1 << 17))
        xs.append(np.ldexp(float(a) ** 3, 3 * int(rng.integers(-300, 300))))
    x = np.array([v for v in xs if np.isfinite(v) and v > 0])
    got = amd.probe_cbrt(x)
    want = np.array([oracle_mod.cbrt_exact(v) for v in x])
    assert_bitwise(got, want, "cbrt near midpoints / exact cubes")


def test_probe_sqrt_div_ieee(amd):
    rng = np.random.default_rng(2)
    a = np.exp(rng.uniform(-300, 300, 50000))
    b = np.exp(rng.uniform(-300, 300, 50000))
    s, q = amd.probe_sqrt_div(a, b)
    assert_bitwise(s, np.sqrt(a), "sqrt")
    assert_bitwise(q, a / b, "div")


def test_probe_range_math(amd):
    # the kernels' bare-FMA sqrt and 1/x against the correctly rounded
    # results over the whole range they are used on: log-uniform inputs,
    # both range ends, exact squares and powers of two
    rng = np.random.default_rng(5)
    lo, hi = 2.0 ** -767, 2.0 ** 1021
    a = np.exp(rng.uniform(np.log(lo), np.log(hi), 200000))
    m = rng.uniform(1.0, 2.0, 20000)           # one binade, densely
    sq = np.floor(rng.uniform(1, 2 ** 26, 10000)) ** 2 * 2.0 ** rng.integers(-700, 900, 10000).astype(float)
    ends = np.array([lo, np.nextafter(lo, 1.0), np.nextafter(hi, 0.0), 1.0, 2.0, 0.5, 4.0, 1e-9, 1e9])
    a = np.concatenate([a, m, sq, ends])
    s, r = amd.probe_range_math(a)
    assert_bitwise(s, np.sqrt(a), "sqrt_pos")
    assert_bitwise(r, 1.0 / a, "rcp_pos")
    s, r = amd.probe_range_math(-a[:1000])
    assert_bitwise(r, 1.0 / -a[:1000], "rcp_pos negative")


def test_probe_fp64_issue_rate(amd):
    # the FP64 yardstick: a sane rate under the no-FMA half of the vector
    # peak (78.6 TFLOP/s counts an FMA as two), a plausible shader clock,
    # and more waves per SIMD never issuing slower
    rates = {}
    for w in (1, 2, 8):
        tf, mhz = amd.probe_fp64(w, 5000)
        assert 1.0 < tf < 39.3 * 1.05, (w, tf)
        assert 500.0 < mhz < 3500.0, (w, mhz)
        rates[w] = tf
    assert rates[8] > rates[1] * 0.95, rates
    tf, mhz = amd.probe_fp64(2, 5000, one_chain=True)
    assert 1.0 < tf < 39.3 * 1.05 and 500.0 < mhz < 3500.0, (tf, mhz)
    with pytest.raises(amd.LuleshError):
        amd.probe_fp64(0)
    with pytest.raises(amd.LuleshError):
        amd.probe_fp64(9)


def _compare_state(g, o, what):
    for f in NODE:
        assert_bitwise(g.field(f), o.field(f), f"{what} {f}")
    for f in ELEM:
        assert_bitwise(g.field(f), o.field(f), f"{what} {f}")


# sizes 1, 2, 3, 5 and 7: a single force tile wider than the mesh,
# single-plane z-chunks, EOS tiles of a few elements
@pytest.mark.parametrize("nx,cycles", [(1, -1), (2, -1), (3, -1), (4, 1), (4, -1), (5, -1),
                                       (7, -1), (8, 5), (8, -1), (16, 40)])
def test_single_domain_bitwise(amd, oracle_mod, nx, cycles):
    o = oracle_mod.Oracle(nx, cbrt_mode=oracle_mod.CBRT_EXACT)
    assert o.run(cycles) == 0
    g = amd.Group(nx)
    if cycles >= 0:
        g.set_max_cycles(cycles)
    assert g.run() == 0
    rc, cyc, t, dt = g.sync()
    assert cyc == o.cycle
    assert t == o.time and dt == o.deltatime
    _compare_state(g, o, f"s={nx} cycles={cycles}")
    g.close()
    o.close()


@pytest.mark.parametrize("nx,cycles", [(20, 30), (45, 12), (80, 6)])
def test_force_tile_shapes_bitwise(amd, oracle_mod, nx, cycles):
    # meshes whose force-march column tiles are not the 18 x 14 of -s 100
    # (22 x 8, 24 x 9, 28 x 9 element slots) and whose z-chunks end in a
    # short chunk: the same state as the oracle, bit for bit
    print(f"case: tile shapes s={nx} {cycles} cycles", flush=True)
    o = oracle_mod.Oracle(nx, cbrt_mode=oracle_mod.CBRT_EXACT)
    assert o.run(cycles) == 0
    g = amd.Group(nx)
    g.set_max_cycles(cycles)
    assert g.run() == 0
    assert g.sync()[1:] == (o.cycle, o.time, o.deltatime)
    _compare_state(g, o, f"s={nx} cycles={cycles}")
    g.close()
    o.close()


def test_eos_volume_clamps_bitwise(amd, oracle_mod):
    # reference volumes scaled so that vnew falls below eosvmin in one
    # element and above eosvmax in another: the EOS clamps (written back
    # into vnew, as the reference does), the eosvmax branches and the
    # compression at both ends of its range, bit for bit
    print("case: EOS volume clamps, s=6, 5 cycles", flush=True)
    o = oracle_mod.Oracle(6, cbrt_mode=oracle_mod.CBRT_EXACT)
    g = amd.Group(6)
    volo = g.field("volo")
    volo[17] *= 1e11
    volo[50] *= 1e-11
    g.domains[0].set_field("volo", volo)
    o.set_field("volo", volo)
    g.set_max_cycles(5)
    assert g.run() == 0 and o.run(5) == 0
    assert g.field("vnew")[17] == 1e-9 and g.field("vnew")[50] == 1e9
    assert g.sync()[1:] == (o.cycle, o.time, o.deltatime)
    _compare_state(g, o, "EOS clamps")


def test_s16_full_run_bitwise(amd, oracle_mod):
    o = oracle_mod.Oracle(16, cbrt_mode=oracle_mod.CBRT_EXACT)
    o.run()
    g = amd.Group(16)
    assert g.run() == 0
    _compare_state(g, o, "s=16 full")
    vg, vo = g.verify(), o.verify()
    for k in vg:
        assert vg[k] == vo[k], k


def test_rep_and_region_invariance(amd):
    # results never depend on -r/-b/-c: the EOS repetitions recompute the
    # same values (LULESH/omp_4.0/lulesh.cc:2677-2815)
    base = amd.Group(8, nr=11, balance=1, cost=1)
    base.run()
    for nr, b, c in [(1, 1, 1), (5, 2, 2), (16, 4, 1)]:
        g = amd.Group(8, nr=nr, balance=b, cost=c)
        g.run()
        assert_bitwise(g.field("e"), base.field("e"), f"r={nr} b={b} c={c}")
        assert_bitwise(g.field("x"), base.field("x"), f"r={nr} b={b} c={c}")


def test_deterministic(amd):
    a = amd.Group(12)
    a.run()
    b = amd.Group(12)
    b.run()
    for f in ["e", "p", "q", "x", "xd"]:
        assert_bitwise(a.field(f), b.field(f), f)


def test_s30_energy_vs_reference(amd, oracle_mod):
    o = oracle_mod.Oracle(30, cbrt_mode=oracle_mod.CBRT_LIBM)
    o.run()
    g = amd.Group(30)
    assert g.run() == 0
    vg, vo = g.verify(), o.verify()
    rel = abs(vg["origin_energy"] - vo["origin_energy"]) / abs(vo["origin_energy"])
    assert rel <= 1e-10, rel
    assert abs(vg["total_abs_diff"] - vo["total_abs_diff"]) <= 1e-8
    assert g.sync()[1] == o.cycle
