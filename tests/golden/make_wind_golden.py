"""Writes worbe2_wind.npz: the wind stress the reference's default EMBM
configuration reads for the worbe2 world (embm-defaults.nml: xu_wstress =
taux_u.interp, yu_wstress = tauy_u.interp, xv_wstress = taux_v.interp,
yv_wstress = tauy_v.interp), one value per line, j outer and i inner
(initialise_embm, embm.f90:845-858), N m-2.

Saved as "stress", (maxj, maxi, 4) = x and y at u points, x and y at v
points, the values parsed from the files' decimal text.  Run from the
repository root with the reference tree present:

    python tests/golden/make_wind_golden.py
"""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/data/embm"
MAXI = MAXJ = 36


def main() -> None:
    st = np.zeros((MAXJ, MAXI, 4))
    for q, name in enumerate(("taux_u", "tauy_u", "taux_v", "tauy_v")):
        v = np.array(open(os.path.join(SRC, name + ".interp")).read().split(), dtype=np.float64)
        st[..., q] = v[:MAXI * MAXJ].reshape(MAXJ, MAXI)
    np.savez_compressed(os.path.join(HERE, "worbe2_wind.npz"), stress=st)
    print("worbe2_wind.npz:", st.shape, "max |stress|", np.abs(st).max())


if __name__ == "__main__":
    main()
