"""Writes carbchem_ref.npz: input/output vectors of the reference's carbonate
chemistry (gem_carbchem.f90: sub_calc_carbconst, the salinity totals,
sub_calc_carb, sub_calc_carb_RF0), from the reference sources compiled in
place into oracle/_ref (``make -C oracle ref``, amdflang -fdefault-real-8).

The vectors pin the CPU restatement (oracle/cgo_oracle.c) where the
reference tree is absent.  Run from the repository root:

    python tests/golden/make_carbchem_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_ffi as O  # noqa: E402

N = 96
CASES = ((1e-3, 100), (1e-12, 1000))  # (carb_tol, carb_maxit)


def main() -> None:
    R = O.ref_carbchem()
    if R is None:
        sys.exit("oracle/_ref is not built: make -C oracle ref (needs the reference tree)")
    rng = np.random.default_rng(2024)
    # inputs: D [m], T [K], S, DIC, ALK, PO4 [mol/kg], starting [H]
    x = np.column_stack([rng.uniform(0, 5500, N), 273.15 + rng.uniform(-3, 38, N), rng.uniform(24, 45, N),
                         rng.uniform(1.8e-3, 2.2e-3, N), rng.uniform(2.2e-3, 2.45e-3, N),
                         rng.uniform(0, 3e-6, N), 10 ** -7.8 * rng.uniform(0.5, 2, N)])
    out = {"inputs": x, "tol_maxit": np.array(CASES, dtype=np.float64)}
    cc_all, tot_all = np.zeros((N, 17)), np.zeros((N, 4))
    for i, (D, T, S, *_rest) in enumerate(x):
        R.ref_calc_carbconst(D, T, S, cc_all[i])
        R.ref_salt_totals(S, tot_all[i])
    out["carbconst"], out["salt_totals"] = cc_all, tot_all
    for c, (tol, maxit) in enumerate(CASES):
        R.ref_set_params(tol, maxit)
        carb, carbalk, status = np.zeros((N, 10)), np.zeros((N, 13)), np.zeros(N)
        rf0 = np.zeros(N)
        for i, (D, T, S, DIC, ALK, PO4, H0) in enumerate(x):
            cc, tot = cc_all[i], tot_all[i]
            c2 = np.zeros(10)
            c2[0] = H0
            status[i] = R.ref_calc_carb(DIC, ALK, tot[0], PO4, 0.0, tot[1], tot[2], tot[3], 0.0, 0.0, cc, c2,
                                        carbalk[i])
            carb[i] = c2
            R.ref_calc_carb_rf0(DIC, ALK, PO4, 0.0, tot[1], tot[2], tot[3], 0.0, 0.0, cc, c2)
            rf0[i] = c2[9]
        out[f"carb_{c}"], out[f"carbalk_{c}"], out[f"status_{c}"], out[f"rf0_{c}"] = carb, carbalk, status, rf0
    path = os.path.join(HERE, "carbchem_ref.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {N} cases x {len(CASES)} tolerances")


if __name__ == "__main__":
    main()
