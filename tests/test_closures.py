"""Closure-table construction (smili_amd.closures) against the reference's
own closure tables of its VLBA 3C273 example (docs/examples/uvdata/
cphase.csv and logcamp.csv, made by VisTable.make_bstable / make_catable,
vistable.py:743-975, 1532-1842), at every fourth timestamp
(tests/golden/closures_3c273.npz).

Those tables were written without the SNR relabelling of stations (they
match ``reorder_st=False`` row for row).  The relabelled path
(``reorder_st=True``, vistable.py:247-293) is pinned by what the closure
algebra fixes: every closure phase is the wrapped phase sum of its
visibilities around the triangle, every log closure amplitude the log of
its amplitude ratio plus the Ei debias, each time tag's set is
independent, and it spans the same space as the un-relabelled set."""
import itertools

import numpy as np
import pandas as pd
import pytest

from smili_amd import closures

FIX = "tests/golden/closures_3c273.npz"
VIS = ("gsthour", "freq", "stokesid", "ifid", "chid", "ch", "u", "v", "w", "st1", "st2", "st1name", "st2name",
       "amp", "phase", "sigma")


@pytest.fixture(scope="module")
def z():
    return np.load(FIX)


@pytest.fixture(scope="module")
def vis(z):
    t = pd.DataFrame({c: z["vis_" + c] for c in VIS})
    t.insert(0, "utc", pd.to_datetime(z["vis_utc"]))
    return t


def _check_rows(got, z, prefix, ints, floats):
    assert len(got) == z[prefix + "_utc"].size
    assert np.array_equal(got["utc"].astype("int64").to_numpy(), z[prefix + "_utc"])
    for c in ints:
        assert np.array_equal(got[c].to_numpy(), z[f"{prefix}_{c}"]), c
    for c in floats:
        ref = z[f"{prefix}_{c}"]
        tol = 1e-13 * np.abs(ref).max()
        if c == "phase":
            d = np.abs((got[c].to_numpy() - ref + 180.0) % 360.0 - 180.0)
        else:
            d = np.abs(got[c].to_numpy() - ref)
        assert d.max() <= tol, (c, d.max())


def test_bstable_matches_reference(vis, z):
    bs = closures.make_bstable(vis, reorder_st=False)
    _check_rows(bs, z, "bs", ("ch", "st1", "st2", "st3", "st1name", "st2name", "st3name", "freq"),
                ("u12", "v12", "w12", "u23", "v23", "w23", "u31", "v31", "w31", "amp", "phase", "sigma"))


def test_catable_matches_reference(vis, z):
    ca = closures.make_catable(vis, reorder_st=False)
    _check_rows(ca, z, "ca", ("ch", "st1", "st2", "st3", "st4", "st1name", "st2name", "st3name", "st4name", "freq"),
                ("u1", "v1", "w1", "u2", "v2", "w2", "u3", "v3", "w3", "u4", "v4", "w4", "amp", "sigma", "logamp",
                 "logsigma"))


def test_tags_split_ifs(vis):
    """Four IFs at 43.0075-43.2155 GHz per timestamp: each is its own tag
    (the 32-bit integer frequencies differ), not one merged tag."""
    t = closures._tagged(vis)
    per_utc = t.groupby("utc")["tag"].nunique()
    nif = vis.groupby("utc")["freq"].nunique()
    assert (per_utc == nif).all() and nif.max() == 4


def test_snr_order(vis):
    """Stations ranked by median |V|/sigma over their baselines, from a
    long-format groupby."""
    snr = vis["amp"] / vis["sigma"]
    long = pd.concat([pd.DataFrame({"st": vis["st1"], "snr": snr}), pd.DataFrame({"st": vis["st2"], "snr": snr})])
    med = long.groupby("st")["snr"].median().sort_values(ascending=False)
    assert closures.check_snr(vis) == med.index.tolist()


def _vis_lookup(t):
    """(utc, ch, name1, name2) -> (amp, phase, u, v) with the conjugate for
    the reversed baseline."""
    d = {}
    for r in t.itertuples(index=False):
        key = (r.utc, r.ch)
        d[key + (r.st1name, r.st2name)] = (r.amp, r.phase, r.u, r.v)
        d[key + (r.st2name, r.st1name)] = (r.amp, -r.phase, -r.u, -r.v)
    return d


def _wrap(x):
    return (np.asarray(x) + 180.0) % 360.0 - 180.0


def _incidence(rows, nst):
    """Closure rows as signed baseline-incidence vectors over station labels
    1..nst."""
    idx = {b: k for k, b in enumerate(itertools.combinations(range(1, nst + 1), 2))}
    out = np.zeros((len(rows), len(idx)))
    for i, cyc in enumerate(rows):
        for (a, b), s in cyc:
            out[i, idx[(min(a, b), max(a, b))]] += s if a < b else -s
    return out


def test_bstable_reordered_closure_relations(vis):
    bs = closures.make_bstable(vis, reorder_st=True)
    ref = closures.make_bstable(vis, reorder_st=False)
    look = _vis_lookup(vis)
    amp, ph, u = [], [], []
    for r in bs.itertuples(index=False):
        k = (r.utc, r.ch)
        a12, a23, a31 = look[k + (r.st1name, r.st2name)], look[k + (r.st2name, r.st3name)], \
            look[k + (r.st3name, r.st1name)]
        amp.append(a12[0] * a23[0] * a31[0])
        ph.append(a12[1] + a23[1] + a31[1])
        u.append((a12[2], a23[2], a31[2]))
    assert np.allclose(bs["amp"], amp, rtol=1e-13, atol=0)
    assert np.abs(_wrap(bs["phase"].to_numpy() - np.asarray(ph))).max() <= 1e-10
    assert np.allclose(bs[["u12", "u23", "u31"]].to_numpy(), np.asarray(u), rtol=1e-13, atol=1e-3)
    # per tag: independent, same count and same span as the un-relabelled set
    names = sorted(set(vis["st1name"]) | set(vis["st2name"]))
    nid = {n: i + 1 for i, n in enumerate(names)}
    for key, g in bs.groupby(["utc", "ch"]):
        g0 = ref[(ref["utc"] == key[0]) & (ref["ch"] == key[1])]
        cyc = lambda df: [(((nid[a], nid[b]), 1), ((nid[b], nid[c]), 1), ((nid[c], nid[a]), 1))  # noqa: E731
                          for a, b, c in zip(df["st1name"], df["st2name"], df["st3name"])]
        M, M0 = _incidence(cyc(g), len(names)), _incidence(cyc(g0), len(names))
        r = np.linalg.matrix_rank(M)
        assert r == len(g) == len(g0)
        assert np.linalg.matrix_rank(np.vstack([M, M0])) == r


def test_catable_reordered_closure_relations(vis):
    from scipy.special import expi
    ca = closures.make_catable(vis, reorder_st=True)
    ref = closures.make_catable(vis, reorder_st=False)
    look = _vis_lookup(vis)
    sig = {}
    for r in vis.itertuples(index=False):
        sig[(r.utc, r.ch, r.st1name, r.st2name)] = sig[(r.utc, r.ch, r.st2name, r.st1name)] = r.sigma
    want = []
    for r in ca.itertuples(index=False):
        k = (r.utc, r.ch)
        pairs = ((r.st1name, r.st2name), (r.st3name, r.st4name), (r.st1name, r.st3name), (r.st2name, r.st4name))
        a = [look[k + p][0] for p in pairs]
        rho = [max(1.0, (x / sig[k + p]) ** 2 - 1) for x, p in zip(a, pairs)]
        bias = 0.5 * (expi(-rho[0] / 2) + expi(-rho[1] / 2) - expi(-rho[2] / 2) - expi(-rho[3] / 2))
        want.append(np.log(a[0] * a[1] / a[2] / a[3]) + bias)
    assert np.allclose(ca["logamp"], want, rtol=0, atol=1e-12)
    names = sorted(set(vis["st1name"]) | set(vis["st2name"]))
    nid = {n: i + 1 for i, n in enumerate(names)}
    for key, g in ca.groupby(["utc", "ch"]):
        g0 = ref[(ref["utc"] == key[0]) & (ref["ch"] == key[1])]
        # amplitudes do not see baseline orientation: unsigned incidence
        M, M0 = _unsigned(g, nid, len(names)), _unsigned(g0, nid, len(names))
        r = np.linalg.matrix_rank(M)
        assert r == len(g) == len(g0)
        assert np.linalg.matrix_rank(np.vstack([M, M0])) == r


def _unsigned(df, nid, nst):
    """log|V12 V34 / V13 V24| rows over baselines: +12 +34 -13 -24."""
    idx = {b: k for k, b in enumerate(itertools.combinations(range(1, nst + 1), 2))}
    out = np.zeros((len(df), len(idx)))
    for i, (a, b, c, d) in enumerate(zip(df["st1name"], df["st2name"], df["st3name"], df["st4name"])):
        for (p, q), s in (((a, b), 1), ((c, d), 1), ((a, c), -1), ((b, d), -1)):
            x, y = nid[p], nid[q]
            out[i, idx[(min(x, y), max(x, y))]] += s
    return out


def _tiny_table(nst=5, utcs=2, drop=()):
    rng = np.random.default_rng(3)
    rows = []
    for ti in range(utcs):
        for a, b in itertools.combinations(range(1, nst + 1), 2):
            if (ti, a, b) in drop:
                continue
            rows.append(dict(utc=pd.Timestamp("2020-01-01") + pd.Timedelta(seconds=10 * ti), gsthour=0.1 * ti,
                             freq=2.3e11, stokesid=1, ifid=0, chid=0, ch=0, u=rng.normal(), v=rng.normal(),
                             w=rng.normal(), st1=a, st2=b, st1name=f"S{a}", st2name=f"S{b}",
                             amp=rng.uniform(1, 2), phase=rng.uniform(-180, 180), sigma=0.1))
    return pd.DataFrame(rows)


def test_counts_and_dependent():
    """Full array of 5 stations: (n-1)(n-2)/2 = 6 independent closure
    phases and n(n-3)/2 = 5 closure amplitudes per tag; dependent=True keeps
    all C(5,3) = 10 triangles.  Without baseline 2-4 at the second timestamp
    the 9 baselines leave E - V + 1 = 5 closure phases and E - V = 4
    closure amplitudes (the graph has a triangle, so it is not bipartite),
    and dependent=True drops the triangles through 2-4."""
    t = _tiny_table(drop={(1, 2, 4)})
    t0 = t["utc"].iloc[0]
    bs = closures.make_bstable(t, reorder_st=False)
    ca = closures.make_catable(t, reorder_st=False)
    assert (bs["utc"] == t0).sum() == 6 and (bs["utc"] != t0).sum() == 5
    assert (ca["utc"] == t0).sum() == 5 and (ca["utc"] != t0).sum() == 4
    dep = closures.make_bstable(t, dependent=True, reorder_st=False)
    assert (dep["utc"] == t0).sum() == 10
    second = dep[dep["utc"] != t0]
    # triangles of 5 stations that avoid baseline 2-4
    assert len(second) == sum(1 for c in itertools.combinations(range(1, 6), 3) if not {2, 4} <= set(c))
    assert all(not {2, 4} <= {a, b, c} for a, b, c in zip(second["st1"], second["st2"], second["st3"]))


def test_redundant_and_empty():
    t = _tiny_table(utcs=1)
    bs = closures.make_bstable(t, redundant=[["S1", "S2"]], dependent=True, reorder_st=False)
    assert len(bs) == sum(1 for c in itertools.combinations(range(1, 6), 3) if not {1, 2} <= set(c))
    assert all(not {1, 2} <= {a, b, c} for a, b, c in zip(bs["st1"], bs["st2"], bs["st3"]))
    two = t[(t["st1"] <= 2) & (t["st2"] <= 2)]
    assert len(closures.make_bstable(two, reorder_st=False)) == 0
    assert len(closures.make_catable(two, reorder_st=False)) == 0
    assert "logamp" in closures.make_catable(two, reorder_st=False).columns
