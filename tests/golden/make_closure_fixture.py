"""Writes tests/golden/closures_3c273.npz: every fourth timestamp of the
reference's example visibility table (docs/examples/uvdata/visibility.csv,
VLBA 3C273 at 43 GHz, four IFs) with the columns the closure builders read,
and the rows of the reference's closure tables made from it
(docs/examples/uvdata/cphase.csv, logcamp.csv) at the same timestamps.
Closures never span timestamps, so the closure rows of a timestamp depend on
that timestamp's visibilities only.  Run in a checkout that has the
reference at /root/reference."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from smili_amd import uvtable  # noqa: E402

SRC = "/root/reference/docs/examples/uvdata"
VIS = ("gsthour", "freq", "stokesid", "ifid", "chid", "ch", "u", "v", "w", "st1", "st2", "st1name", "st2name",
       "amp", "phase", "sigma")
BS = ("freq", "ch", "st1", "st2", "st3", "st1name", "st2name", "st3name", "u12", "v12", "w12", "u23", "v23", "w23",
      "u31", "v31", "w31", "amp", "phase", "sigma")
CA = ("freq", "ch", "st1", "st2", "st3", "st4", "st1name", "st2name", "st3name", "st4name", "u1", "v1", "w1",
      "u2", "v2", "w2", "u3", "v3", "w3", "u4", "v4", "w4", "amp", "sigma", "logamp", "logsigma")


def _cols(t, prefix, cols):
    out = {f"{prefix}_utc": t["utc"].astype("int64").to_numpy()}
    for c in cols:
        x = t[c].to_numpy()
        out[f"{prefix}_{c}"] = x.astype(str) if x.dtype == object else x
    return out


def main():
    vis = uvtable.read_vistable(os.path.join(SRC, "visibility.csv"))
    keep = sorted(vis["utc"].unique())[::4]
    out = {}
    for prefix, fname, reader, cols in (("vis", "visibility.csv", uvtable.read_vistable, VIS),
                                        ("bs", "cphase.csv", uvtable.read_bstable, BS),
                                        ("ca", "logcamp.csv", uvtable.read_catable, CA)):
        t = reader(os.path.join(SRC, fname))
        t = t[t["utc"].isin(keep)].reset_index(drop=True)
        out.update(_cols(t, prefix, cols))
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "closures_3c273.npz"), **out)
    print({k: v.shape for k, v in out.items() if k.endswith("_utc")})


if __name__ == "__main__":
    main()
