/*
 * clmsoil_oracle.h — CPU restatement of CLM.jl's per-column soil solves.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the parity oracle and the CPU
 * baseline. Only the tests, the smoke check and the benchmark's CPU-baseline leg
 * may load it; the product library (libclmsoil) never calls or links it.
 *
 * Every function follows the statement order of the CLM.jl source it cites
 * (paths relative to the CLM.jl tree) using the same argument structs as the
 * product C ABI (include/clmsoil.h), with host pointers instead of device ones.
 * It is compiled with -ffp-contract=off so that, like Julia, a*b+c rounds twice.
 */
#ifndef CLMSOIL_ORACLE_H
#define CLMSOIL_ORACLE_H

#include "../include/clmsoil.h"
#include "../include/clmsoil_zd09.h"
#include "../include/clmsoil_ghf.h"
#include "../include/clmsoil_hydro.h"
#include "../include/clmsoil_surfhyd.h"
#include "../include/clmsoil_lake.h"
#include "../include/clmsoil_lvt.h"
#include "../include/clmsoil_ch4.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Julia Base.max / Base.min on Float64 (NaN-propagating, +0 > -0). */
double orc_jl_max(double x, double y);
double orc_jl_min(double x, double y);

/* Clapp-Hornberger SWRC, swrc_clapp_hornberg.jl:43-55, 76-88, 109-117. */
void   orc_soil_hk(double hksat, double bsw, double s, double imped,
                   double *hk, double *dhkds);
void   orc_soil_suction(double bsw, double sucsat, double s,
                        double *smp, double *dsmpds);
double orc_soil_suction_inverse(double bsw, double sucsat, double smp_target);
/* ice_impedance, soil_water_movement.jl:722-724 */
double orc_ice_impedance(double icefrac, double e_ice);

/* Scalar helpers, soil_water_movement.jl:736-773, 1080-1140 (1-D vectors, 0-based
 * storage of the reference's 1-based levels). */
void orc_tridiagonal_col(double *u, const double *a, const double *b,
                         const double *c, const double *r, int jtop, int lbj, int ubj);
void orc_compute_rhs_moisture_form(int nlayers, const double *vert_trans_sink,
                                   const double *qin, const double *qout,
                                   const double *dt_dz, double *rmx);
void orc_compute_lhs_moisture_form(int nlayers, const double *dt_dz,
                                   const double *dqidw0, const double *dqidw1,
                                   const double *dqodw1, const double *dqodw2,
                                   double *amx, double *bmx, double *cmx);

/* tridiagonal_solve!, tridiagonal.jl:35-63 (single column, no pivot guard). */
void orc_tridiagonal_solve(double *u, const double *a, const double *b,
                           const double *c, const double *r, int jtop, int nlevs);

/* Batched operators: same signatures as the C ABI (host memory). */
int orc_tridiagonal_multi(double *u, const double *a, const double *b,
                          const double *c, const double *r, const int64_t *jtop,
                          const uint8_t *mask, int64_t ncols, int64_t nlevs,
                          int64_t ld);
int orc_batched_band_solve(double *tvector, const double *bmatrix,
                           const double *rvector, const int64_t *jtop,
                           const int64_t *jbot, const uint8_t *mask, int32_t kl,
                           int32_t ku, int32_t nlevsno_off, int64_t nc,
                           int64_t nlev, int64_t ld);

/* Moisture-form Richards step, soil_water_movement.jl:1250-1517.
 * dbg_qin/dbg_qout (optional, [ld, nlevsoi]) receive the last iteration's
 * interface fluxes, for the telescoping identity qin[j] == qout[j-1]. */
int orc_soilwm_moisture_form_solve(const clmsoil_mf_args *a,
                                   const clmsoil_levels *lev, double dtime,
                                   const clmsoil_swm_config *cfg,
                                   double *dbg_qin, double *dbg_qout);
/* soilwater_moisture_form!, soil_water_movement.jl:1532-1601. */
int orc_soilwater_moisture_form(const clmsoil_mf_args *a, const clmsoil_levels *lev,
                                double dtime, const clmsoil_swm_config *cfg);

/* Non-urban soil_temperature!, soil_temperature.jl:358-593. */
int orc_soil_temperature(const clmsoil_heat_args *a, const clmsoil_levels *lev,
                         double dtime, const clmsoil_heat_config *cfg);
/* Stages of one heat column (0-based column c), as the reference's tests call
 * them. orc_heat_assemble_probe runs soil_therm_prop! through the solve and
 * fn1 (t_soisno of column c is updated) and copies out the interface
 * conductivity tk and heat capacity cv [nlevsno+nlevgrnd], the band rows
 * bmatrix[(b-1) + 5*(row-1)] and rvector [nlevsno+1+nlevgrnd], and tk_h2osfc.
 * orc_phase_change_h2osfc / orc_phase_change_beta run phase_change_h2osfc!
 * (:1757-1892) / phase_change_beta! (:1902-2289) alone, with the given
 * c_h2osfc and top-snow fact / the column's fact [nlevsno+nlevgrnd]. */
int orc_heat_assemble_probe(const clmsoil_heat_args *a, const clmsoil_levels *lev,
                            double dtime, const clmsoil_heat_config *cfg, int64_t c,
                            double *tk, double *cv, double *bmatrix, double *rvector,
                            double *tk_h2osfc);
int orc_phase_change_h2osfc(const clmsoil_heat_args *a, const clmsoil_levels *lev,
                            double dtime, int64_t c, double c_h2osfc, double fact_top);
int orc_phase_change_beta(const clmsoil_heat_args *a, const clmsoil_levels *lev, double dtime,
                          int64_t c, const double *fact);

/* Julia's x^4 for Float64 (Base.Math.pow_body: compensated power by squaring). */
double orc_jl_pow4(double x);
/* Ground heat flux, compute_ground_heat_flux_and_deriv! (soil_temperature.jl:1200-1263),
 * in ghf_oracle.c. */
int orc_ground_heat_flux(const clmsoil_ghf_args *a, const clmsoil_levels *lev);
/* Soil energy balance error, soil_fluxes.jl:380-447, 499-525 (ghf_oracle.c). */
int orc_soil_energy_balance(const clmsoil_seb_args *a, const clmsoil_levels *lev,
                            double dtime);
/* Soil water fractions and water table, soil_hydrology.jl:110-136, 676-893
 * (hydro_oracle.c). */
int orc_set_soil_water_fractions(const clmsoil_fractions_args *a, const clmsoil_levels *lev);
int orc_water_table(const clmsoil_water_table_args *a, const clmsoil_levels *lev, double dtime,
                    double aq_sp_yield_min, int recompute_frost_table);
/* Water tables from the moisture profile, soil_hydrology.jl:1522-1715 and
 * src/infrastructure/kernels.jl:209-218 */
int orc_perched_water_table(const clmsoil_theta_wt_args *a, const clmsoil_levels *lev);
int orc_theta_based_water_table(const clmsoil_theta_wt_args *a, const clmsoil_levels *lev);
int orc_clamp_zwt_to_bedrock(const clmsoil_theta_wt_args *a, const clmsoil_levels *lev);

/* Surface hydrology before the Richards step, clm_driver.jl:1939-2012
 * (surfhyd_oracle.c). scratch2nc: 2*nc doubles for update_h2osfc!'s arrays. */
void orc_sh_floodc(const clmsoil_surfhyd_args *a);
void orc_sh_sat_excess(const clmsoil_surfhyd_args *a, const clmsoil_surfhyd_params *p);
void orc_sh_qflx_inputs(const clmsoil_surfhyd_args *a);
void orc_sh_infil_excess(const clmsoil_surfhyd_args *a, const clmsoil_surfhyd_params *p);
void orc_sh_route(const clmsoil_surfhyd_args *a, const clmsoil_surfhyd_params *p);
void orc_sh_update_h2osfc(const clmsoil_surfhyd_args *a, const clmsoil_surfhyd_params *p,
                          double dtime, double *surf, double *drain);
void orc_sh_infiltration(const clmsoil_surfhyd_args *a);
void orc_sh_surf_runoff(const clmsoil_surfhyd_args *a, const clmsoil_surfhyd_params *p,
                        int nlevsno, double dtime);
void orc_sh_urban_pond(const clmsoil_surfhyd_args *a, const clmsoil_surfhyd_params *p,
                       int nlevsno, double dtime);
int orc_surface_hydrology(const clmsoil_surfhyd_args *a, const clmsoil_levels *lev, double dtime,
                          const clmsoil_surfhyd_params *p, double *scratch2nc);
void orc_set_surfhyd_jitter(uint64_t seed);
int orc_compute_effec_rootfrac_and_vert_tran_sink(const clmsoil_rootsink_args *a,
                                                  const clmsoil_levels *lev);

/* Lake temperature solve, lake_temperature.jl:632-785, 1205-1225 (lake_oracle.c);
 * dbg NULL or 11*N*nc doubles of the assembly's scratch arrays. */
int orc_lake_temperature_solve(const clmsoil_lake_solve_args *a, const clmsoil_levels *lev,
                               double dtime, double *dbg);

/* Soil organic matter vertical transport, litter_vert_transp.jl:93-639 (lvt_oracle.c) */
double orc_jl_pow5(double x);
int orc_litter_vert_transp(const clmsoil_lvt_args *a, double dtime, const clmsoil_lvt_params *p);

/* Methane and oxygen transport, methane.jl:101-129, 1663-2092 (ch4_oracle.c).
 * orc_set_ch4_jitter moves every exp/log/pow result one ulp (seed 0: off);
 * orc_ch4_debug sets the copy-outs of the next calls (see ch4_oracle.c). */
int orc_ch4_tran(const clmsoil_ch4_tran_args *a, double dtime, const clmsoil_ch4_tran_params *p);
void orc_set_ch4_jitter(uint64_t seed);
void orc_ch4_debug(double *err_raw, double *rows, double *khcc);
double orc_ch4_jl_pow(double x, double y);
/* the transport's diffusivity of species s (0 CH4, 1 O2) at one level; gas != 0
 * at or above the water table (ch4_oracle.c, :1818-1840) */
double orc_ch4_diffusivity(const clmsoil_ch4_tran_params *p, int s, int gas, double t, double ws,
                           double vol_min, double cellorg, double bsw, double liq, double ice);
/* Test hook: a non-zero seed moves every pow() of the Zeng-Decker step one ulp
 * (direction from a hash of the operands and the seed); 0 restores exact pow(). */
void orc_set_zd_pow_jitter(uint64_t seed);
/* Zeng-Decker 2009 step, soilwater_zengdecker2009! (soil_water_movement.jl:1616-1764). */
int orc_soilwater_zengdecker2009(const clmsoil_zd_args *a, const clmsoil_levels *lev,
                                 double dtime, const clmsoil_swm_config *cfg);

/* op 0: clm_pow(x, y), 1: clm_log(x), 2: clm_log10(x), 3: clm_exp(x),
 * 4: Julia max(x, y), 5: Julia min(x, y), and the fast paths on their domains:
 * 6: clm_pow_pos(x, y), 7: clm_pow_base(clm_log_ext(x), y) (x the base),
 * 8: clm_log10_pos(x) */
void orc_clm_math(int op, const double *x, const double *y, double *out, int64_t n);
/* Number of OpenMP threads the batched functions use. */
int orc_num_threads(void);
/* Test hook of the moisture-form step: each substep iteration of column `col`
 * (0-based; -1 = off) appends (dtsub, errorMax, accepted) to buf[3*k..3*k+2],
 * at most `cap` rows; orc_mf_trace_count returns how many were written. */
void orc_set_mf_trace(int64_t col, double *buf, int cap);
int orc_mf_trace_count(void);

#ifdef __cplusplus
}
#endif
#endif
