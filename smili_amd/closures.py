"""Closure-quantity tables from a visibility table, with the selection rules
of SMILI's VisTable methods (smili/uvdata/uvtable/vistable.py):

* ``uvsort`` (86-115): baselines with st1 > st2 are swapped (u, v and phase
  change sign; w does not) and the table is sorted by utc, st1, st2,
  stokesid, ch;
* ``station_dic`` (295-309), ``check_snr`` / ``reorder_st_snr`` (247-293):
  stations relabelled 1..N in decreasing median SNR;
* ``make_bstable`` (743-975): closure phases over the triangles of each time
  tag (utc, stokesid, integer frequency), an independent set chosen in
  lexicographic station order by matrix rank (``calc_matrix_bs``,
  3546-3565) unless ``dependent``; bispectrum amplitude, phase
  phi12 + phi23 - phi13 wrapped to (-180, 180], sigma from the relative
  errors;
* ``make_catable`` (1532-1842): log closure amplitudes over the quadrangles
  of each tag, up to three per quadrangle in the reference's order, the
  same rank rule (``calc_matrix_ca``, 3523-3544), log-amplitude errors from
  the SNRs and the Ei-function debiasing.

Each time tag is grouped once with numpy instead of row-by-row table
lookups; the selection (which triangles / quadrangles, in which order, with
which station labels) follows the reference step by step.
"""
from __future__ import annotations

import itertools

import numpy as np


def _pd():
    import pandas as pd
    return pd


def station_dic(table, id2name=True):
    """vistable.py:295-309."""
    st1 = table.drop_duplicates(subset="st1")
    st2 = table.drop_duplicates(subset="st2")
    if id2name:
        out = dict(zip(st1.st1.values, st1.st1name.values))
        out.update(dict(zip(st2.st2.values, st2.st2name.values)))
    else:
        out = dict(zip(st1.st1name.values, st1.st1.values))
        out.update(dict(zip(st2.st2name.values, st2.st2.values)))
    return out


def uvsort(table):
    """vistable.py:86-115."""
    out = table.copy()
    t = (out["st1"] > out["st2"]).to_numpy()
    for a, b in (("st1", "st2"), ("st1name", "st2name")):
        c1 = out.loc[t, b].values.copy()
        c2 = out.loc[t, a].values.copy()
        out.loc[t, b] = c2
        out.loc[t, a] = c1
    out.loc[t, ["u", "v", "phase"]] *= -1
    return out.sort_values(by=["utc", "st1", "st2", "stokesid", "ch"])


def check_snr(table, avgtype="median"):
    """Station ids in decreasing median (or mean) SNR (vistable.py:247-273)."""
    pd = _pd()
    dic = station_dic(table)
    names, ids = list(dic.values()), list(dic.keys())
    snr = (table["amp"] / table["sigma"]).to_numpy()
    s1 = table["st1name"].to_numpy()
    s2 = table["st2name"].to_numpy()
    vals = []
    for nm in names:
        sel = snr[(s1 == nm) | (s2 == nm)]
        vals.append(np.median(sel) if avgtype == "median" else np.mean(sel))
    t = pd.DataFrame({"stname": names, "st": ids, "snr": vals}).sort_values("snr", ascending=False)
    return t["st"].to_list()


def reorder_st_snr(table, avgtype="median"):
    """vistable.py:275-293."""
    order = check_snr(table, avgtype)
    out = table.copy()
    for new, old in enumerate(order, start=1):
        for col in ("st1", "st2"):
            out.loc[table[col] == old, col] = new
    return uvsort(out)


def _getblid(st1, st2, nst):
    """vistable.py:3487-3503."""
    a, b = min(st1, st2), max(st1, st2)
    return a * nst - a * (a + 1) // 2 + b - a - 1


def _nontrivial(baselines, redundant):
    """vistable.py:3506-3514."""
    if redundant is None:
        return True
    for bl, red in itertools.product(baselines, redundant):
        if bl[0] in red and bl[1] in red:
            return False
    return True


class _Rank:
    """Rows +-1 on baseline ids; accepts a row when it raises the rank
    (np.linalg.matrix_rank of the stacked rows, calc_matrix_* 3523-3565)."""

    def __init__(self, nbl, dependent):
        self.rows = []
        self.nbl = nbl
        self.dependent = dependent
        self.rank = 0

    def try_add(self, plus, minus):
        row = np.zeros(self.nbl)
        row[list(plus)] = 1
        row[list(minus)] = -1
        if self.dependent:
            self.rows.append(row)
            self.rank = 1
            return True
        r = int(np.linalg.matrix_rank(np.asarray(self.rows + [row])))
        if r > self.rank:
            self.rows.append(row)
            self.rank = r
            return True
        return False


def _canon_redundant(table, redundant):
    if redundant is None:
        return None
    n2i = station_dic(table, id2name=False)
    out = []
    for grp in redundant:
        ids = [n2i[s] for s in grp if isinstance(s, str) and s in n2i] + [s for s in grp if not isinstance(s, str)]
        out.append(sorted(set(ids)))
    return out


def _tagged(table):
    """Sorted table with bl and tag columns (vistable.py:791-810): a new tag
    where utc, stokesid or the (32-bit) integer frequency changes."""
    t = table.sort_values(by=["utc", "stokesid", "ch", "st1", "st2"]).reset_index(drop=True)
    t["bl"] = t["st1"] * 256 + t["st2"]
    utc = t["utc"].to_numpy()
    sid = t["stokesid"].to_numpy()
    # the reference compares np.int32(freq): frequencies in Hz wrap to 32
    # bits, which keeps IFs apart (their differences are far below 2^32 Hz)
    fq = (t["freq"].to_numpy().astype(np.int64) & 0xFFFFFFFF).astype(np.uint32).view(np.int32)
    new = np.ones(len(t), bool)
    new[1:] = (utc[1:] != utc[:-1]) | (sid[1:] != sid[:-1]) | (fq[1:] != fq[:-1])
    new[0] = False
    t["tag"] = np.cumsum(new)
    return t


def _groups(t):
    tags = t["tag"].to_numpy()
    starts = np.concatenate([[0], np.nonzero(np.diff(tags))[0] + 1, [len(t)]])
    return [(starts[i], starts[i + 1]) for i in range(len(starts) - 1)]


def _first_rows(t, a, b):
    """First row of each baseline inside rows [a, b) (the reference takes
    bltab.loc[0] of the tag's rows of that baseline)."""
    out = {}
    bls = t["bl"].to_numpy()
    for k in range(a, b):
        out.setdefault(int(bls[k]), k)
    return out


def make_bstable(table, redundant=None, dependent=False, reorder_st=True):
    """Closure phases (bispectra) of a visibility table (vistable.py:743-975)."""
    pd = _pd()
    vt = reorder_st_snr(table) if reorder_st else table.copy()
    stdict = station_dic(vt)
    redundant = _canon_redundant(vt, redundant)
    t = _tagged(vt)
    cols = {k: t[k].to_numpy() for k in ("u", "v", "w", "amp", "phase", "sigma", "st1", "st2")}
    blsets, combos = [], []
    keys = ("utc,gsthour,freq,stokesid,ifid,chid,ch,st1,st2,st3,st1name,st2name,st3name,"
            "u12,v12,w12,u23,v23,w23,u31,v31,w31,amp,phase,sigma").split(",")
    out = {k: [] for k in keys}
    for a, b in _groups(t):
        bl = t["bl"].to_numpy()[a:b]
        blset = sorted(set(int(x) for x in bl))
        stset = sorted(set(int(x) for x in np.concatenate([cols["st1"][a:b], cols["st2"][a:b]])))
        if len(blset) < 3 or len(stset) < 3:
            continue
        key = tuple(blset)
        if key in blsets:
            cmb = combos[blsets.index(key)]
        else:
            nst = len(stset)
            nblmax = nst * (nst - 1) // 2
            nbsmax = (nst - 1) * (nst - 2) // 2
            rk = _Rank(nblmax, dependent)
            cmb = []
            bs = set(blset)
            for i1, i2, i3 in itertools.combinations(range(nst), 3):
                if rk.rank >= nbsmax and not dependent:
                    break
                s1, s2, s3 = stset[i1], stset[i2], stset[i3]
                b12, b23, b13 = s1 * 256 + s2, s2 * 256 + s3, s1 * 256 + s3
                if not _nontrivial([[s1, s2], [s1, s3], [s2, s3]], redundant):
                    continue
                if not (b12 in bs and b23 in bs and b13 in bs):
                    continue
                if rk.try_add((_getblid(i1, i2, nst), _getblid(i2, i3, nst)), (_getblid(i1, i3, nst),)):
                    cmb.append(((b12, b23, b13), (s1, s2, s3)))
            if rk.rank == 0:
                cmb = None
            blsets.append(key)
            combos.append(cmb)
        if not cmb:
            continue
        first = _first_rows(t, a, b)
        for (b12, b23, b13), (s1, s2, s3) in cmb:
            k1, k2, k3 = first[b12], first[b23], first[b13]
            r1 = cols["sigma"][k1] / cols["amp"][k1]
            r2 = cols["sigma"][k2] / cols["amp"][k2]
            r3 = cols["sigma"][k3] / cols["amp"][k3]
            amp = cols["amp"][k1] * cols["amp"][k2] * cols["amp"][k3]
            for k in ("utc", "gsthour", "freq", "stokesid", "ifid", "chid", "ch"):
                out[k].append(t[k].iat[a])
            for k, s in (("st1", s1), ("st2", s2), ("st3", s3)):
                out[k].append(s)
                out[k + "name"].append(stdict[s])
            out["u12"].append(cols["u"][k1]); out["v12"].append(cols["v"][k1]); out["w12"].append(cols["w"][k1])
            out["u23"].append(cols["u"][k2]); out["v23"].append(cols["v"][k2]); out["w23"].append(cols["w"][k2])
            out["u31"].append(-cols["u"][k3]); out["v31"].append(-cols["v"][k3]); out["w31"].append(-cols["w"][k3])
            out["amp"].append(amp)
            out["phase"].append(cols["phase"][k1] + cols["phase"][k2] - cols["phase"][k3])
            out["sigma"].append(amp * np.sqrt(r1 ** 2 + r2 ** 2 + r3 ** 2))
    o = pd.DataFrame(out)
    for a_, b_ in (("12", "12"), ("23", "23"), ("31", "31")):
        o["uvdist" + a_] = np.sqrt(np.square(o["u" + b_]) + np.square(o["v" + b_]))
    d = np.asarray([o["uvdist12"], o["uvdist23"], o["uvdist31"]])
    o["uvdistave"] = d.mean(axis=0) if d.size else []
    o["uvdistmax"] = d.max(axis=0) if d.size else []
    o["uvdistmin"] = d.min(axis=0) if d.size else []
    ph = np.deg2rad(o["phase"].to_numpy(np.float64))
    o["phase"] = np.rad2deg(np.arctan2(np.sin(ph), np.cos(ph)))
    cols_out = ["utc", "gsthour", "freq", "stokesid", "chid", "ifid", "ch",
                "u12", "v12", "w12", "uvdist12", "u23", "v23", "w23", "uvdist23",
                "u31", "v31", "w31", "uvdist31", "uvdistmin", "uvdistmax", "uvdistave",
                "st1", "st2", "st3", "st1name", "st2name", "st3name", "amp", "phase", "sigma"]
    return o[cols_out].reset_index(drop=True)


def make_catable(table, redundant=None, dependent=False, debias=True, reorder_st=True):
    """Log closure amplitudes of a visibility table (vistable.py:1532-1842)."""
    from scipy.special import expi
    pd = _pd()
    vt = reorder_st_snr(table) if reorder_st else table.copy()
    stdict = station_dic(vt)
    redundant = _canon_redundant(vt, redundant)
    t = _tagged(vt)
    cols = {k: t[k].to_numpy() for k in ("u", "v", "w", "amp", "sigma", "st1", "st2")}
    blsets, combos = [], []
    keys = ("utc,gsthour,freq,stokesid,ifid,chid,ch,st1,st2,st3,st4,st1name,st2name,st3name,st4name,"
            "u1,v1,w1,u2,v2,w2,u3,v3,w3,u4,v4,w4,amp,sigma,logamp,logsigma").split(",")
    out = {k: [] for k in keys}
    for a, b in _groups(t):
        bl = t["bl"].to_numpy()[a:b]
        blset = sorted(set(int(x) for x in bl))
        stset = sorted(set(int(x) for x in np.concatenate([cols["st1"][a:b], cols["st2"][a:b]])))
        if len(blset) < 4 or len(stset) < 4:
            continue
        key = tuple(blset)
        if key in blsets:
            cmb = combos[blsets.index(key)]
        else:
            nst = len(stset)
            nblmax = nst * (nst - 1) // 2
            ncamax = nst * (nst - 3) // 2
            rk = _Rank(nblmax, dependent)
            cmb = []
            bs = set(blset)
            for i1, i2, i3, i4 in itertools.combinations(range(nst), 4):
                s1, s2, s3, s4 = stset[i1], stset[i2], stset[i3], stset[i4]
                b12, b13, b14 = s1 * 256 + s2, s1 * 256 + s3, s1 * 256 + s4
                b23, b24, b34 = s2 * 256 + s3, s2 * 256 + s4, s3 * 256 + s4
                g = lambda x, y: _getblid(x, y, nst)  # noqa: E731
                ncomb = 0
                # the three candidate ratios of a quadrangle, in the
                # reference's order, each with its own rank test
                cands = (
                    ([[s1, s4], [s2, s3]], (b12, b34, b13, b24), (s1, s2, s3, s4),
                     (g(i1, i2), g(i3, i4)), (g(i1, i3), g(i2, i4))),
                    ([[s1, s2], [s3, s4]], (b13, b24, b14, b23), (s1, s3, s4, s2),
                     (g(i1, i3), g(i2, i4)), (g(i1, i4), g(i2, i3))),
                    ([[s1, s3], [s2, s4]], (b12, b34, b14, b23), (s1, s2, s4, s3),
                     (g(i1, i2), g(i3, i4)), (g(i1, i4), g(i2, i3))),
                )
                stop = False
                for ci, (nt, bls, sts, plus, minus) in enumerate(cands):
                    if ci == 2 and ncomb > 1:
                        break
                    if rk.rank >= ncamax and not dependent:
                        stop = True
                        break
                    if not _nontrivial(nt, redundant):
                        continue
                    if not all(x in bs for x in bls):
                        continue
                    if rk.try_add(plus, minus):
                        cmb.append((bls, sts))
                        if ci < 2:
                            ncomb += 1
                if stop:
                    break
            if rk.rank == 0:
                cmb = None
            blsets.append(key)
            combos.append(cmb)
        if not cmb:
            continue
        first = _first_rows(t, a, b)
        for bls, sts in cmb:
            ks = [first[x] for x in bls]
            rho = [np.nanmax([1, np.square(cols["amp"][k] / cols["sigma"][k]) - 1]) for k in ks]
            amp = cols["amp"][ks[0]] * cols["amp"][ks[1]] / cols["amp"][ks[2]] / cols["amp"][ks[3]]
            logamp = np.log(amp)
            logsigma = np.sqrt(rho[0] ** -1 + rho[1] ** -1 + rho[2] ** -1 + rho[3] ** -1)
            sigma = amp * logsigma
            if debias:
                bias = expi(-rho[0] / 2.)
                bias += expi(-rho[1] / 2.)
                bias -= expi(-rho[2] / 2.)
                bias -= expi(-rho[3] / 2.)
                bias *= 0.5
                logamp += bias
                amp *= np.exp(bias)
            for k in ("utc", "gsthour", "freq", "stokesid", "ifid", "chid", "ch"):
                out[k].append(t[k].iat[a])
            for n, s in enumerate(sts, start=1):
                out[f"st{n}"].append(s)
                out[f"st{n}name"].append(stdict[s])
            for n, k in enumerate(ks, start=1):
                out[f"u{n}"].append(cols["u"][k]); out[f"v{n}"].append(cols["v"][k]); out[f"w{n}"].append(cols["w"][k])
            out["amp"].append(amp)
            out["sigma"].append(sigma)
            out["logamp"].append(logamp)
            out["logsigma"].append(logsigma)
    o = pd.DataFrame(out)
    for n in range(1, 5):
        o[f"uvdist{n}"] = np.sqrt(np.square(o[f"u{n}"]) + np.square(o[f"v{n}"]))
    d = np.asarray([o[f"uvdist{n}"] for n in range(1, 5)])
    o["uvdistave"] = d.mean(axis=0) if d.size else []
    o["uvdistmax"] = d.max(axis=0) if d.size else []
    o["uvdistmin"] = d.min(axis=0) if d.size else []
    cols_out = ["utc", "gsthour", "freq", "stokesid", "ifid", "chid", "ch",
                "u1", "v1", "w1", "uvdist1", "u2", "v2", "w2", "uvdist2",
                "u3", "v3", "w3", "uvdist3", "u4", "v4", "w4", "uvdist4",
                "uvdistmin", "uvdistmax", "uvdistave",
                "st1", "st2", "st3", "st4", "st1name", "st2name", "st3name", "st4name",
                "amp", "sigma", "logamp", "logsigma"]
    return o[cols_out].reset_index(drop=True)
