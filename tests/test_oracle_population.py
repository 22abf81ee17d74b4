"""Pins the oracle's population routines: injection laws (KS against the
analytic CDFs), removal and split invariants, and the distribution
diagnostics against numpy histograms."""
import numpy as np
import pytest
from scipy import stats

import oracle
from helpers import ESC_DIST, cfg2d, escaped_particles, face_layout, random_states

BOX = (0.1, 0.2, 0.0, 0.6, 0.9, 1.0)


def test_inject_delta_uniform_box():
    c = cfg2d(32)
    prm = oracle.params(c)
    n = 5000
    out, tag_max = oracle.inject(prm, np.zeros(0, oracle.PARTICLE_DTYPE), 10000, n, 1e-7, 1, 5.0,
                                 0.5, 0.25, BOX, 7.0)
    assert len(out) == n and tag_max == n
    np.testing.assert_array_equal(out["tag_injected"], np.arange(n))
    assert np.all(out["tag_splitted"] == 1) and np.all(out["count_flag"] == 1)
    assert np.all(out["p"] == c["p0"]) and np.all(out["weight"] == 1.0)
    assert np.all(out["v"] == 5.0 * out["p"] / c["p0"])
    assert np.all(out["dt"] == 1e-7)
    mu_max = float(np.float32(0.99))
    assert np.all(np.abs(out["mu"]) <= mu_max)
    for key, lo, hi in (("x", BOX[0], BOX[3]), ("y", BOX[1], BOX[4]), ("z", BOX[2], BOX[5]),
                        ("t", 0.5, 0.75)):
        u = (out[key] - lo) / (hi - lo)
        assert stats.kstest(u, "uniform").pvalue > 1e-3, key


def test_inject_power_law_and_maxwellian_laws():
    c = cfg2d(32, pmin=1e-2, pmax=10.0, p0=0.1)
    prm = oracle.params(c)
    n = 6000
    # power law f(p) ~ p^-q on [p0, pmax] by inversion
    q = 3.0
    out, _ = oracle.inject(prm, np.zeros(0, oracle.PARTICLE_DTYPE), n, n, 1e-7, 2, 1.0, 0.0, 1.0,
                           BOX, q)
    p0, pm = c["p0"], c["pmax"]

    def cdf_pl(p):
        return (p ** (1 - q) - p0 ** (1 - q)) / (pm ** (1 - q) - p0 ** (1 - q))

    assert stats.kstest(out["p"], cdf_pl).pvalue > 1e-3
    # Maxwellian rejection sampling of x^2 exp(-x^2), x = p/p0 on [pmin, pmax]/p0
    out, _ = oracle.inject(prm, np.zeros(0, oracle.PARTICLE_DTYPE), n, n, 1e-7, 0, 1.0, 0.0, 1.0,
                           BOX, q)
    xs = np.linspace(c["pmin"] / p0, pm / p0, 200001)
    pdf = xs ** 2 * np.exp(-xs ** 2)
    cdf = np.concatenate([[0], np.cumsum(0.5 * (pdf[1:] + pdf[:-1]) * np.diff(xs))])
    cdf /= cdf[-1]
    assert stats.kstest(out["p"] / p0, lambda v: np.interp(v, xs, cdf)).pvalue > 1e-3


def test_inject_past_capacity_overwrites_last_slot():
    prm = oracle.params(cfg2d(16))
    out, tag_max = oracle.inject(prm, np.zeros(0, oracle.PARTICLE_DTYPE), 5, 8, 1e-7, 1, 1.0, 0.0,
                                 1.0, BOX, 7.0)
    assert len(out) == 5 and tag_max == 8
    np.testing.assert_array_equal(out["tag_injected"], [0, 1, 2, 3, 7])


def test_remove_particles_invariants():
    c = cfg2d(16)
    a = random_states(c, 1000, seed=4)
    rng = np.random.default_rng(4)
    a["count_flag"] = rng.choice([1, 1, 1, 0, -1, -2, -3, -4], size=len(a))
    kept, nesc = oracle.remove(a)
    assert np.all(kept["count_flag"] == 1)
    assert nesc == int(np.sum(a["count_flag"] < 0))
    assert sorted(kept["tag_injected"]) == sorted(a["tag_injected"][a["count_flag"] == 1])
    # the survivors carry their state unchanged
    order = np.argsort(kept["tag_injected"])
    src = a[a["count_flag"] == 1]
    np.testing.assert_array_equal(kept[order]["x"], src[np.argsort(src["tag_injected"])]["x"])


def test_split_particle_invariants():
    c = cfg2d(16, p0=1.0, pmax=1e4)
    prm = oracle.params(c)
    a = random_states(c, 2000, seed=8)
    a["p"] = np.exp(np.random.default_rng(8).uniform(np.log(0.3), np.log(40.0), len(a)))
    rng = np.random.default_rng(9)
    k = rng.integers(0, 3, len(a)).astype(np.int8)
    a["split_times"] = k
    a["weight"] = 0.5 ** k
    a["tag_splitted"] = 1
    ratio, pmin_split = 2.72, 2.0
    thr = pmin_split * 1.0 * ratio ** k
    cand = (a["p"] > thr) & (a["p"] <= c["pmax"])
    out, nsplit = oracle.split(prm, a, 10000, ratio, pmin_split)
    assert nsplit == cand.sum() and len(out) == len(a) + nsplit
    # total weight conserved; children follow in candidate order
    assert out["weight"].sum() == a["weight"].sum()
    children = out[len(a):]
    np.testing.assert_array_equal(children["tag_injected"], a["tag_injected"][cand])
    np.testing.assert_array_equal(children["tag_splitted"], 1 + 2 ** k[cand])
    np.testing.assert_array_equal(children["split_times"], k[cand] + 1)
    np.testing.assert_array_equal(out["split_times"][: len(a)][cand], k[cand] + 1)
    np.testing.assert_array_equal(children["p"], a["p"][cand])
    # (tag_injected, tag_splitted) stays a unique particle id
    ids = set(zip(out["tag_injected"].tolist(), out["tag_splitted"].tolist()))
    assert len(ids) == len(out)
    # at capacity, only the first candidates in array order split
    room = 17
    out2, ns2 = oracle.split(prm, a, len(a) + room, ratio, pmin_split)
    assert ns2 == room and len(out2) == len(a) + room
    first = np.flatnonzero(cand)[:room]
    np.testing.assert_array_equal(out2[len(a):]["tag_injected"], a["tag_injected"][first])
    assert np.all(out2["split_times"][np.flatnonzero(cand)[room:]] == k[cand][room:])


def _dist_cfg(map_bins=8, npbins=16):
    return dict(npp_global=64, nmu_global=1, local_dist=1,
                local=[dict(enabled=1, npbins=npbins, nmu=1, rx=4, ry=4, rz=1, pmin=1e-2, pmax=1e4),
                       None, None,
                       dict(enabled=1, npbins=4, nmu=2, rx=8, ry=2, rz=1, pmin=0.1, pmax=100.0)])


def test_distributions_match_numpy_histograms():
    c = cfg2d(32)
    prm = oracle.params(c)
    a = random_states(c, 20000, seed=12, slack=False)
    rng = np.random.default_rng(12)
    a["p"] = np.exp(rng.uniform(np.log(5e-3), np.log(2e4), len(a)))
    a["weight"] = 0.5 ** rng.integers(0, 4, len(a))
    a["mu"] = rng.uniform(-1, 1, len(a))
    dcfg = _dist_cfg()
    fg, locs = oracle.distributions(prm, dcfg, a)
    edges = 10.0 ** np.linspace(-2, 4, 65)
    sel = (a["p"] > 1e-2) & (a["p"] <= 1e4)
    h, _ = np.histogram(a["p"][sel], bins=edges, weights=a["weight"][sel])
    np.testing.assert_allclose(fg, h, rtol=0, atol=1e-12)
    # local set 1: (ip, ix, iy) with ip < npbins; top momentum bin dropped
    f1 = locs[0].reshape(1, 8, 8, 16, 1)[0]  # (iy, ix, ip, imu)
    pe = 10.0 ** np.linspace(-2, 4, 17)
    H, _ = np.histogramdd(np.stack([a["y"], a["x"], a["p"]], 1),
                          bins=[np.linspace(0, 1, 9), np.linspace(0, 1, 9), pe], weights=a["weight"])
    H[..., -1] = 0.0
    np.testing.assert_allclose(f1[..., 0], H, atol=1e-12)
    # local set 4 with two mu bins
    f4 = locs[3].reshape(16, 4, 4, 2)  # (iy, ix, ip, imu)
    pe4 = 10.0 ** np.linspace(-1, 2, 5)
    H4, _ = np.histogramdd(np.stack([a["y"], a["x"], a["p"], a["mu"]], 1),
                           bins=[np.linspace(0, 1, 17), np.linspace(0, 1, 5), pe4,
                                 np.linspace(-1, 1, 3)], weights=a["weight"])
    H4[:, :, -1, :] = 0.0
    np.testing.assert_allclose(f4, H4, atol=1e-12)


def test_quick_check_matches_numpy():
    c = cfg2d(16)
    a = random_states(c, 3000, seed=13)
    a["weight"] = 0.5 ** np.random.default_rng(1).integers(0, 3, len(a))
    q = oracle.quick_check(a, 7, 1.5, 0.25)
    assert q["nptl_current"] == len(a) and q["nptl_split"] == 7
    assert q["ntot"] == a["weight"].sum()
    assert q["leak"] == 1.5 and q["leak_negp"] == 0.25
    assert q["pdt_min"] == min(1.0, a["dt"].min()) and q["pdt_max"] == a["dt"].max()
    assert q["pdt_avg"] == pytest.approx(a["dt"].mean(), rel=1e-13)
    assert q["pmax"] == a["p"].max()


def split_faces(c, dcfg, k, flat):
    layout, _ = face_layout(c, dcfg, k)
    faces, off = {}, 0
    for name, shp in layout:
        m = int(np.prod(shp))
        faces[name] = flat[off:off + m].reshape(shp)
        off += m
    assert off == len(flat)
    return faces


@pytest.mark.parametrize("ndim", [2, 3])
def test_escaped_distributions_match_numpy_histograms(ndim):
    from helpers import cfg3d
    c = cfg2d(32) if ndim == 2 else cfg3d(12)
    prm = oracle.params(c)
    a = escaped_particles(c, 30000, seed=20 + ndim)
    others = random_states(c, 500, seed=5)
    others["count_flag"] = 0    # removed but not escaped: never binned
    fe, locs = oracle.escaped_distributions(prm, ESC_DIST, np.concatenate([a, others]))
    fe = fe.reshape(2 * ndim, 64, 2)
    pe = 10.0 ** np.linspace(-2, 4, 65)
    for f in range(1, 2 * ndim + 1):
        s = a[(a["count_flag"] == -f) & (a["p"] > 1e-2) & (a["p"] <= 1e4)]
        H, _, _ = np.histogram2d(s["p"], s["mu"], bins=[pe, [-1.0, 0.0, 1.0]], weights=s["weight"])
        np.testing.assert_allclose(fe[f - 1], H, rtol=0, atol=1e-12, err_msg=f"face {f}")
    lo = np.array([c["xmin"], c["ymin"], c["zmin"]])
    ln = np.array([c["lx"], c["ly"], c["lz"]])
    names = "xyz"
    for k in (0, 3):
        L = ESC_DIST["local"][k]
        faces = split_faces(c, ESC_DIST, k, locs[k])
        _, nr = face_layout(c, ESC_DIST, k)
        pe_k = 10.0 ** np.linspace(np.log10(L["pmin"]), np.log10(L["pmax"]), L["npbins"] + 1)
        mue = np.linspace(-1.0, 1.0, L["nmu"] + 1)
        for axis in range(ndim):
            b_ax, c_ax = [q for q in range(3) if q != axis]
            for side in (0, 1):
                s = a[a["count_flag"] == -(2 * axis + side + 1)]
                coords = np.stack([s["xyz"[c_ax]], s["xyz"[b_ax]], s["p"], s["mu"]], 1)
                H, _ = np.histogramdd(coords, bins=[
                    np.linspace(lo[c_ax], lo[c_ax] + ln[c_ax], nr[c_ax] + 1),
                    np.linspace(lo[b_ax], lo[b_ax] + ln[b_ax], nr[b_ax] + 1), pe_k, mue],
                    weights=s["weight"])
                H[:, :, -1, :] = 0.0     # the top momentum bin is dropped
                np.testing.assert_allclose(faces[names[axis]][side], H, rtol=0, atol=1e-12,
                                           err_msg=f"set {k + 1} axis {axis} side {side}")


def test_escaped_weight_equals_leak_through_open_faces():
    """Every particle the mover flags as escaped lands in the escaped
    distributions, and their weight is the leak the boundary pass counted."""
    from helpers import nodes_for
    c = cfg2d(32, pbcx=1, pbcy=1, dt_max_rel=1e-2)
    prm = oracle.params(c)
    fa = oracle.gradients(prm, nodes_for(c))
    box = (0.0, 0.0, 0.0, 1.0, 1.0, 0.0)
    a, _ = oracle.inject(prm, np.zeros(0, oracle.PARTICLE_DTYPE), 3000, 3000, 1e-7, 1, 5.0, 0.0,
                         0.2, box, 7.0)
    kept, acc, removed = oracle.mover(prm, fa, a, 0.0, 0.2, nthreads=4, return_removed=True)
    esc = removed[removed["count_flag"] < 0]
    assert len(esc) == acc.nptl_escaped > 20
    assert len(kept) + len(removed) == len(a)
    fe, _ = oracle.escaped_distributions(prm, dict(npp_global=256, nmu_global=1), removed)
    assert fe.sum() == acc.leak
    per_face = fe.reshape(4, -1).sum(1)
    for f in range(1, 5):
        assert per_face[f - 1] == esc["weight"][esc["count_flag"] == -f].sum()


# ---------------------------------------------------------------------------
# Threshold injections and their cell counters

def node_quantities(c, fa, quantity):
    """numpy restatement of the counters' node quantity over interior nodes
    (C order (z, y, x)); |jz| and |J| in FP32 as the Fortran's real(sp) farray."""
    nd = c["ndim"]
    sl = (slice(2, -2) if nd == 3 else slice(0, 1), slice(2, -2), slice(2, -2))
    if quantity == "divv":
        # the counter's array is rebound to start at 1: node (i, j, k) reads
        # (i-2, j-2, k-2) of the ghosted array (no z ghosts in 2D)
        sl = (slice(0, -4) if nd == 3 else slice(0, 1), slice(0, -4), slice(0, -4))
        d = fa[sl + (8,)].astype(np.float64) + fa[sl + (12,)].astype(np.float64)
        if nd == 3:
            d = d + fa[sl + (16,)].astype(np.float64)
        return -d
    f = fa[sl]
    if quantity == "jz":
        return np.abs(f[..., 23] - f[..., 21]).astype(np.float64)
    if quantity == "absj":
        a, b, cc = f[..., 25] - f[..., 27], f[..., 26] - f[..., 22], f[..., 21] - f[..., 23]
        return np.sqrt(a * a + b * b + cc * cc).astype(np.float64)
    return f[..., 3].astype(np.float64)


def in_box_nodes(c, box):
    nd = c["ndim"]
    x = c["xmin"] + np.arange(c["nx"]) * c["dx"]
    y = c["ymin"] + np.arange(c["ny"]) * c["dy"]
    z = c["zmin"] + np.arange(c["nz"]) * c["dz"]
    ix = (x > box[0]) & (x < box[3])
    iy = (y > box[1]) & (y < box[4])
    iz = (z > box[2]) & (z < box[5]) if nd == 3 else np.ones(1, bool)
    return iz[:, None, None] & iy[None, :, None] & ix[None, None, :]


def interp_component(c, fa, comp, x, y, z):
    """Bilinear / trilinear interpolation of one farray component (numpy)."""
    nd = c["ndim"]
    px, py = (x - c["xmin"]) / c["dx"], (y - c["ymin"]) / c["dy"]
    i, j = np.floor(px).astype(int) + 2, np.floor(py).astype(int) + 2
    rx, ry = px - np.floor(px), py - np.floor(py)
    if nd == 2:
        g = fa[0, :, :, comp].astype(np.float64)
        return (g[j, i] * (1 - rx) * (1 - ry) + g[j, i + 1] * rx * (1 - ry) +
                g[j + 1, i] * (1 - rx) * ry + g[j + 1, i + 1] * rx * ry)
    pz = (z - c["zmin"]) / c["dz"]
    k, rz = np.floor(pz).astype(int) + 2, pz - np.floor(pz)
    g = fa[..., comp].astype(np.float64)
    out = 0.0
    for dk, wk in ((0, 1 - rz), (1, rz)):
        for dj, wj in ((0, 1 - ry), (1, ry)):
            for di, wi in ((0, 1 - rx), (1, rx)):
                out = out + g[k + dk, j + dj, i + di] * wk * wj * wi
    return out


def point_quantities(c, fa, quantity, x, y, z):
    q = lambda k: interp_component(c, fa, k, x, y, z)
    if quantity == "jz":
        return np.abs(q(23) - q(21))
    if quantity == "absj":
        return np.sqrt((q(25) - q(27)) ** 2 + (q(26) - q(22)) ** 2 + (q(21) - q(23)) ** 2)
    if quantity == "divv":
        d = q(8) + q(12) + (q(16) if c["ndim"] == 3 else 0.0)
        return -d
    return q(3)


def threshold_case(ndim):
    from helpers import cfg3d, nodes_for
    c = cfg2d(32) if ndim == 2 else cfg3d(12)
    nodes = nodes_for(c)
    # a density that varies, so that the rho injection has something to select
    zz, yy, xx = np.meshgrid(*[np.arange(s) for s in nodes.shape[:3]], indexing="ij")
    nodes[..., 3] = (1.0 + 0.5 * np.sin(0.3 * xx + 0.2 * yy + 0.1 * zz)).astype(np.float32)
    prm = oracle.params(c)
    return c, nodes, prm, oracle.gradients(prm, nodes)


THRESHOLDS = {"jz": 3.0, "absj": 3.0, "divv": 0.3, "rho": 1.2}
TBOX = (0.1, 0.2, 0.05, 0.85, 0.9, 0.95)


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("quantity", ["jz", "absj", "divv", "rho"])
def test_threshold_cell_counts_match_numpy(ndim, quantity):
    c, _, prm, fa = threshold_case(ndim)
    thr = THRESHOLDS[quantity]
    n = oracle.ncells_above_threshold(prm, fa, quantity, thr, TBOX)
    ref = int(np.sum((node_quantities(c, fa, quantity) > thr) & in_box_nodes(c, TBOX)))
    assert n == ref and 0 < n < in_box_nodes(c, TBOX).sum()


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("quantity", ["jz", "absj", "divv", "rho"])
def test_threshold_injection_places_particles_above_threshold(ndim, quantity):
    c, _, prm, fa = threshold_case(ndim)
    thr = THRESHOLDS[quantity]
    empty = np.zeros(0, oracle.PARTICLE_DTYPE)
    a, tag_max, ninj = oracle.inject_threshold(prm, fa, empty, 5000, quantity, 2000, 1e-7, 1, 5.0,
                                               0.0, 0.1, True, thr, 1, TBOX, 7.0)
    assert ninj == len(a) == 2000 and tag_max == 2000
    for k, lo, hi in (("x", 0, 3), ("y", 1, 4), ("z", 2, 5)):
        assert np.all((a[k] >= TBOX[lo]) & (a[k] <= TBOX[hi]))
    v = point_quantities(c, fa, quantity, a["x"], a["y"], a["z"])
    assert np.all(v >= thr * (1 - 1e-12))
    # the accepted points cover the region where the quantity passes: the
    # fraction of uniform box points that pass matches the density of hits
    rng = np.random.default_rng(3)
    u = [rng.uniform(TBOX[i], TBOX[i + 3], 20000) for i in range(3)]
    if ndim == 2:
        u[2] = np.full(20000, 0.5)
    frac = np.mean(point_quantities(c, fa, quantity, *u) >= thr)
    assert 0.0 < frac < 1.0
    # with ncells_norm, nptl_inject = int(nptl * ncells / ncells_norm)
    ncells = oracle.ncells_above_threshold(prm, fa, quantity, thr, TBOX)
    _, _, n2 = oracle.inject_threshold(prm, fa, empty, 5000, quantity, 1000, 1e-7, 1, 5.0, 0.0,
                                       0.1, False, thr, 4 * ncells, TBOX, 7.0)
    assert n2 == 250


def test_threshold_injection_edge_cases():
    c, _, prm, fa = threshold_case(2)
    empty = np.zeros(0, oracle.PARTICLE_DTYPE)
    # no node passes: nothing is injected
    _, _, n = oracle.inject_threshold(prm, fa, empty, 100, "jz", 50, 1e-7, 1, 5.0, 0.0, 0.1, True,
                                      1e9, 1, TBOX, 7.0)
    assert n == 0
    # a threshold the loop's initial value already passes (rho: 0 >= -1): no
    # point is drawn and every particle sits at the box's low corner
    a, _, n = oracle.inject_threshold(prm, fa, empty, 100, "rho", 20, 1e-7, 1, 5.0, 0.0, 0.1, True,
                                      -1.0, 1, TBOX, 7.0)
    assert n == 20
    assert np.all(a["x"] == TBOX[0]) and np.all(a["y"] == TBOX[1]) and np.all(a["z"] == TBOX[2])
