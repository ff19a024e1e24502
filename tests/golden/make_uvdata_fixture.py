"""Writes tests/golden/uvdata_3c273.npz: the columns the imager reads from
the reference's example tables (docs/examples/uvdata/visibility.csv,
cphase.csv, logcamp.csv: VLBA 3C273, 43 GHz), every other timestamp, read
with smili_amd.uvtable.read_*.  Run in a checkout that has the reference at
/root/reference."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from smili_amd import uvtable  # noqa: E402

SRC = "/root/reference/docs/examples/uvdata"
COLS = {
    "vis": ("visibility.csv", uvtable.read_vistable, ("u", "v", "amp", "phase", "sigma")),
    "bs": ("cphase.csv", uvtable.read_bstable, ("u12", "v12", "u23", "v23", "u31", "v31", "amp", "phase", "sigma")),
    "ca": ("logcamp.csv", uvtable.read_catable, ("u1", "v1", "u2", "v2", "u3", "v3", "u4", "v4", "logamp",
                                                   "logsigma")),
}


def main():
    out = {}
    vis = uvtable.read_vistable(os.path.join(SRC, "visibility.csv"))
    keep = sorted(vis["utc"].unique())[::2]
    for key, (fname, reader, cols) in COLS.items():
        t = reader(os.path.join(SRC, fname))
        t = t[t["utc"].isin(keep)].reset_index(drop=True)
        for c in cols:
            out[f"{key}_{c}"] = t[c].to_numpy(np.float64)
        out[f"{key}_utc"] = t["utc"].astype("int64").to_numpy()
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "uvdata_3c273.npz"), **out)
    print({k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
