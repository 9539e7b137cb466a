"""The oracle's pins catch planted mistakes.

Each mutant below is a copy of oracle/cgo_oracle.c with one plausible
mistake planted at a cited line of the restatement (a wrong metric index,
swapped upwind weights, a dropped term or factor).  It is compiled on the
CPU into a scratch library, every oracle call is routed to it, and the pin
tests of tests/test_oracle.py run against it: at least one must fail.  A
mutant whose text is not found fails this test too, so the list cannot go
stale silently when the oracle changes.
"""
import concurrent.futures as cf
import os
import subprocess

import pytest

import oracle_ffi as O
import test_oracle as T
import test_surflux as S
import test_velocity as V

SRC = os.path.join(O.ORACLE_DIR, "cgo_oracle.c")

# (name, text in cgo_oracle.c, replacement, occurrences)
MUTANTS = [
    # tstepo_flux, goldstein.f90:2525-2558
    ("meridional diffusion metric cv2(j-1)",
     "fn[l - 1] = fn[l - 1] - g->cv2[j] *", "fn[l - 1] = fn[l - 1] - g->cv2[j - 1] *", 1),
    ("meridional upwind weights swapped",
     "((1.0 - ups[2]) * TS1(c, l, i, j + 1, k) + (1.0 + ups[2]) * TS1(c, l, i, j, k))",
     "((1.0 + ups[2]) * TS1(c, l, i, j + 1, k) + (1.0 - ups[2]) * TS1(c, l, i, j, k))", 1),
    ("vertical upwind weights swapped",
     "((1.0 - ups[3]) * TS1(c, l, i, j, k + 1) + (1.0 + ups[3]) * TS1(c, l, i, j, k))",
     "((1.0 + ups[3]) * TS1(c, l, i, j, k + 1) + (1.0 - ups[3]) * TS1(c, l, i, j, k))", 1),
    ("zonal upwind weights swapped",
     "((1.0 - ups[1]) * TS1(c, l, i + 1, j, k) + (1.0 + ups[1]) * TS1(c, l, i, j, k))",
     "((1.0 + ups[1]) * TS1(c, l, i + 1, j, k) + (1.0 - ups[1]) * TS1(c, l, i, j, k))", 1),
    ("zonal diffusion metric rc instead of rc2",
     "(TS1(c, l, i + 1, j, k) - TS1(c, l, i, j, k)) * g->rc2[j] * diff1",
     "(TS1(c, l, i + 1, j, k) - TS1(c, l, i, j, k)) * g->rc[j] * diff1", 1),
    ("vertical diffusion metric rdz instead of rdza",
     "(TS1(c, l, i, j, k + 1) - TS1(c, l, i, j, k)) * g->rdza[k] * diffv",
     "(TS1(c, l, i, j, k + 1) - TS1(c, l, i, j, k)) * g->rdz[k] * diffv", 1),
    ("meridional update metric rdsv instead of rds",
     "(fn[l - 1] - FS(l, i)) * g->rds[j]", "(fn[l - 1] - FS(l, i)) * g->rdsv[j]", 1),
    ("isoneutral factor 2 dropped",
     "tv = tv + (2 * dzrho * DXTS(l, ina)", "tv = tv + (dzrho * DXTS(l, ina)", 1),
    # co, goldstein.f90:2732-2737
    ("convective mix not thickness weighted",
     "sum[l - 1] = TS(c, l, i, j, kk[m]) * dzm[kk[m]];", "sum[l - 1] = TS(c, l, i, j, kk[m]) * dzm[kk[m - 1]];", 1),
    # sub_box_remin_part, biogem_box.f90:2555-2596, 2779-2792; gem_util.f90:79-98
    ("CaCO3 dissolution returns 1 ALK",
     "cv->c[OG_S_CACO3][3] = 2.0;", "cv->c[OG_S_CACO3][3] = 1.0;", 1),
    ("CaCO3 e-folding length from POC",
     "double f1 = (1.0 - exp(-dD / p->CaCO3_eL1));", "double f1 = (1.0 - exp(-dD / p->POC_eL1));", 1),
    ("layer thickness ratio inverted",
     "const double layerratio = st->dD[kk + 1] / st->dD[kk];",
     "const double layerratio = st->dD[kk] / st->dD[kk + 1];", 1),
    ("POC O2 demand without the POP share",
     "p->red_POP_PO2 / p->red_POP_POC - cO2POP / p->red_POP_POC", "p->red_POP_PO2 / p->red_POP_POC", 1),
    # biogem_tracercoupling, biogem.f90:1975-2061
    ("coupling de-normalisation without the local salinity",
     "go_ts[O3(st, l, i, j, k)] * st->ocn[O3(st, 1, i, j, k)] * rSold;", "go_ts[O3(st, l, i, j, k)] * rSold;", 1),
    ("coupling re-normalisation dropped",
     "(Snew / st->ocn[O3(st, 1, i, j, k)]) * st->ocn[O3(st, l, i, j, k)];", "st->ocn[O3(st, l, i, j, k)];", 1),
    ("coupling salinity ratio on the tracers dropped",
     "st->ocn[O3(st, l, i, j, k)] = Sratio * v;", "st->ocn[O3(st, l, i, j, k)] = v;", 1),
    ("cell masses scaled by Sratio instead of 1/Sratio",
     "st->M[C3(st, i, j, k)] = rSratio * st->M[C3(st, i, j, k)];",
     "st->M[C3(st, i, j, k)] = Sratio * st->M[C3(st, i, j, k)];", 1),
    # air-sea exchange, biogem_box.f90:81-299, gem_carbchem.f90:1367-1426; atchem.f90:63-158
    ("Schmidt number cubic term with the wrong sign",
     "c[2] * TC2 - c[3] * TC3", "c[2] * TC2 + c[3] * TC3", 1),
    ("O2 solubility without the density",
     "return bunsen ? x / (rho * CONST_V) : x;", "return bunsen ? x / CONST_V : x;", 1),
    ("flux limiter ratio inverted",
     "fo = (AIRSEA_R_DFLUX_DEQM_MAX / r) * fo;", "fo = r * fo;", 1),
    ("air-sea flux enters the ocean with the wrong sign",
     "*r = *r + dtrM * (0.0 - st->fatm[2 * c]);", "*r = *r + dtrM * st->fatm[2 * c];", 1),
    ("atmosphere mixed without the gas temperature",
     "st->atm[2 * c + q] = (tot / totV) * CONV_PA_ATM * CONST_R_SI * CONST_ZEROC;",
     "st->atm[2 * c + q] = (tot / totV) * CONV_PA_ATM * CONST_R_SI;", 1),
    # the velocity solve, goldstein.f90:2845-3714, goldstein_lib.f90:186-241
    ("psi equation: periodic west neighbour in the wrong column",
     "if (i == 1) l = 2 * n + 1;", "if (i == MI) l = 2 * n + 1;", 1),
    ("barotropic u without the cos(latitude) metric",
     "UBL(ubv, 1, i, j) = -MRH(1, i, j) * g->c[j] * (PSIL(psv, i, j)",
     "UBL(ubv, 1, i, j) = -MRH(1, i, j) * (PSIL(psv, i, j)", 1),
    ("baroclinic shear: drag and Coriolis swapped",
     "dzu[0][k] = -(g->s[j] * tv1 + MDRAG(1, i, j) * tv2) * MRTV(i, j);",
     "dzu[0][k] = -(MDRAG(1, i, j) * tv1 + g->s[j] * tv2) * MRTV(i, j);", 1),
    ("depth mean of the baroclinic velocity not removed",
     "MU(u, 1, i, j, k) = MU(u, 1, i, j, k) - sum[0] * MRH(1, i, j) + UBL(mo->ub, 1, i, j);",
     "MU(u, 1, i, j, k) = MU(u, 1, i, j, k) + UBL(mo->ub, 1, i, j);", 1),
    ("island circulation: Coriolis term with the wrong sign",
     "cor = -g->s[jpi] * 0.25 *", "cor = g->s[jpi] * 0.25 *", 1),
    ("continuity: meridional divergence with rdsv",
     "MU(u, 2, i, j - 1, k) * g->cv[j - 1]) * g->rds[j];", "MU(u, 2, i, j - 1, k) * g->cv[j - 1]) * g->rdsv[j];", 1),
    ("bottom pressure integrated without the layer spacing",
     "(MRHO(i, j, k) + MRHO(i, j, k - 1)) * g->dza[k - 1] * 0.5;", "(MRHO(i, j, k) + MRHO(i, j, k - 1)) * 0.5;", 1),
    ("drag: equatorial band one row narrower",
     "if (kloc2 > kmxdrg || abs(j - MJ / 2) <= jeb)", "if (kloc2 > kmxdrg || abs(j - MJ / 2) < jeb)", 1),
    # EMBM surflux, embm.f90:2756-3679, and the surface slot, goldstein.f90:149-170
    ("ocean heat loss: sensible heat with the wrong sign",
     "const double fx0o = (1 - ca) * fxsw - fxsen - fxlw", "const double fx0o = (1 - ca) * fxsw + fxsen - fxlw", 1),
    ("atmosphere heat gain: longwave with the wrong sign",
     "const double fx0oa = ca * fxsw + fxlata + fxsen + fxlw - fxplw;",
     "const double fx0oa = ca * fxsw + fxlata + fxsen - fxlw - fxplw;", 1),
    ("runoff destination indices swapped",
     "F2(runoff, ir, jr) = F2(runoff, ir, jr) + F2(pptn, i, j);",
     "F2(runoff, jr, ir) = F2(runoff, jr, ir) + F2(pptn, i, j);", 1),
    ("runoff area weight from the source row",
     "F2(pptn, i, j) * g->ds[j] * g->rds[jr];", "F2(pptn, i, j) * g->ds[j] * g->rds[j];", 1),
    ("Dalton number bounds swapped",
     "ce = fmaxf_(6.0E-5, fminf_(2.19E-3, ce));", "ce = fminf_(6.0E-5, fmaxf_(2.19E-3, ce));", 1),
    ("ice surface: conduction with the wrong sign",
     "E_CONSIC * (tsfreez - ticold);", "E_CONSIC * (ticold - tsfreez);", 1),
    ("ice Newton step: derivative of the emission with the wrong sign",
     "const double dtieq = sich * (-4.0 * E_EMO", "const double dtieq = sich * (4.0 * E_EMO", 1),
    ("CO2 forcing with the wrong sign",
     "- p->delf2x * log(cco2 / E_CO20)", "+ p->delf2x * log(cco2 / E_CO20)", 1),
    ("insolation month one step late",
     "const int imth = (int)((istot - 1) % p->nyear) + 1;", "const int imth = (int)(istot % p->nyear) + 1;", 1),
    ("wind speed: one polar row zonally averaged",
     "if (j <= 2 || j >= maxj - 1)", "if (j <= 1 || j >= maxj)", 1),
    ("rain without the humidity ceiling rmax",
     "(*ashum - p->rmax * qsata) * E_RHOAO", "(*ashum - qsata) * E_RHOAO", 1),
    ("freshwater adjustment dropped",
     "O(OG_SF_PPTN_OCN, i, j) + F2(p->pmeadj, i, j);", "O(OG_SF_PPTN_OCN, i, j);", 1),
    ("freezing point: S instead of sqrt(S)",
     "(-0.0575 + 0.0017 * sqrt(salt) - 0.0002 * salt)", "(-0.0575 + 0.0017 * salt - 0.0002 * salt)", 1),
    ("ocean latent heat without the ice formed in open water",
     "(-(E_RHO0 * E_HLV * F2(evap, i, j)) + fmaxf_(0.0, qb - fx0o))", "(-(E_RHO0 * E_HLV * F2(evap, i, j)))", 1),
    ("heat slot with the wrong sign",
     "sflux[0 + (size_t)maxl * c] = -fx0neto * rfluxsc;", "sflux[0 + (size_t)maxl * c] = fx0neto * rfluxsc;", 1),
    ("freshwater slot left in mm",
     "fwfxneto = fwfxneto * E_MM2M;", "fwfxneto = fwfxneto;", 1),
    ("wind speed: row 1 without the half v stress",
     "const double tv2 = j == 1 ? F2(tau2, i, j) / 2", "const double tv2 = j == 1 ? F2(tau2, i, j)", 1),
    ("sea-ice albedo rising with air temperature",
     "0.40 - 0.04 * atemp", "0.40 + 0.04 * atemp", 1),
    ("ice growth cap: open-water growth kept",
     "O(OG_SF_DTHSIC, i, j) = sica * dhsic + (1 - sica) * dho;", "O(OG_SF_DTHSIC, i, j) = sica * dhsic + dho;", 1),
]

# the pins (tests/test_oracle.py, tests/test_velocity.py, tests/test_surflux.py), cheapest first;
# the carbonate chemistry's compiled-reference comparison is represented by
# its golden vectors (the first case of test_carbchem_matches_reference_vectors)
PINS = [
    T.test_eos_polynomial, T.test_grid_closed_forms, T.test_carbconst_closed_form_mehrbach,
    T.test_carbonate_root_satisfies_alkalinity, T.test_convection_mixes_inversion_to_thickness_weighted_mean,
    T.test_vertical_diffusion_matches_textbook_column, T.test_vertical_upwind_advection_empties_one_cell_at_courant_one,
    T.test_zonal_upwind_advection_shifts_one_cell_at_courant_one, T.test_zonal_diffusion_discrete_eigenmode,
    T.test_meridional_upwind_advection_shifts_one_row_at_courant_one,
    T.test_meridional_diffusion_sin_latitude_eigenvector, T.test_uniform_tracer_stays_uniform,
    T.test_oracle_random_velocity_conserves, T.test_isoneutral_flux_closed_form,
    T.test_biogem_oracle_uptake_closed_form, T.test_dom_remin_closed_form, T.test_particulate_remin_closed_form,
    T.test_caco3_remin_closed_form, T.test_coupling_identity_and_single_cell_anomaly,
    T.test_coupling_uniform_salinity_anomaly, lambda: T.test_biogem_oracle_conserves_phosphorus("config2"),
    lambda: T.test_carbchem_matches_reference_vectors(0), T.test_oracle_conserves_inventory,
    T.test_airsea_solubility_closed_forms, T.test_airsea_piston_velocity_wanninkhof,
    T.test_airsea_bulk_flux_equilibrium_and_limit, T.test_airsea_step_conserves_and_atmosphere_mixes,
    V.test_channel_barotropic_flow_is_the_friction_wind_balance, V.test_channel_ekman_layer_and_depth_mean,
    V.test_worbe2_island_circulation_vanishes, V.test_worbe2_solve_is_linear_in_the_forcing,
    V.test_worbe2_velocity_leaves_a_uniform_tracer_uniform, V.test_worbe2_no_wind_and_level_density_give_no_flow,
    V.test_drag_values, V.test_bottom_pressure_is_the_hydrostatic_integral,
    S.test_usurf_inverts_the_bulk_stress_formula, S.test_usurf_polar_rows_are_zonal_means,
    S.test_ocean_and_atmosphere_heat_closures, S.test_air_sea_exchange_cancels_between_air_and_sea,
    S.test_rain_is_the_humidity_the_clamp_removes, S.test_saturation_humidity_bolton,
    S.test_freezing_point_millero, S.test_runoff_arrives_where_routed_and_totals_the_land_rain,
    S.test_freshwater_adjustment_and_evaporation_signs, S.test_bulk_transfer_coefficient_is_observed_magnitude,
    S.test_sea_ice_albedo_falls_with_air_temperature, S.test_sea_ice_temperature_is_the_surface_energy_balance_root,
    S.test_sea_ice_growth_cap, S.test_co2_doubling_forcing, S.test_compound_co2_rise, S.test_seasonal_cycle_period,
    S.test_surface_slot_heats_and_freshens_the_top_layer,
]
# the reference's own island() where it is built (oracle/_ref): the one pin of
# the island integral's Coriolis term, which no closed form fixes (the
# integral round an island is not path-independent in this discretisation)
if O.ref_gold() is not None:
    PINS += [V.test_island_integral_matches_the_reference_worbe2,
             V.test_island_integral_matches_the_reference_u_and_v_path]
NEEDS_REF = {"island circulation: Coriolis term with the wrong sign"}


def _build(tmp, idx, name, find, repl, count):
    src = open(SRC).read()
    n = src.count(find)
    if n != count:
        return idx, None, f"{name}: text found {n} times, expected {count}"
    d = os.path.join(tmp, f"m{idx}")
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "cgo_oracle.c"), "w") as f:
        f.write(src.replace(find, repl))
    with open(os.path.join(O.ORACLE_DIR, "cgo_oracle.h")) as h, open(os.path.join(d, "cgo_oracle.h"), "w") as f:
        f.write(h.read())
    so = os.path.join(d, "libmut.so")
    r = subprocess.run(["gcc", "-O2", "-fPIC", "-std=gnu11", "-ffp-contract=off", "-shared", "-o", so,
                        os.path.join(d, "cgo_oracle.c"), "-lm"], capture_output=True, text=True)
    if r.returncode != 0:
        return idx, None, f"{name}: gcc failed: {r.stderr[-400:]}"
    return idx, so, None


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("mutants"))
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        res = list(ex.map(lambda a: _build(tmp, *a), [(i, *m) for i, m in enumerate(MUTANTS)]))
    return {i: (so, err) for i, so, err in res}


def test_unmutated_oracle_passes_every_pin():
    for pin in PINS:
        pin()


@pytest.mark.parametrize("idx", range(len(MUTANTS)), ids=[m[0] for m in MUTANTS])
def test_pins_catch_planted_mistake(mutant_libs, idx):
    so, err = mutant_libs[idx]
    assert err is None, err
    if MUTANTS[idx][0] in NEEDS_REF and O.ref_gold() is None:
        pytest.skip("pinned only by the reference build (oracle/_ref)")
    caught = None
    with O.use_library(so):
        for pin in PINS:
            try:
                pin()
            except Exception as ex:  # an AssertionError, or the pin's own arithmetic failing on bad output
                caught = (getattr(pin, "__name__", "pin"), type(ex).__name__)
                break
    assert caught is not None, f"no pin fails with the mistake planted: {MUTANTS[idx][0]}"
