// Derived in part from Xassemblianist/WFE, Copyright (c) 2026 Omer Kaan (Xassemblianist),
// MIT License; the licence text is in THIRD_PARTY_NOTICES at the repository root.
/*
 * oracle.h -- CPU FP64 restatement of the WFE large-step dynamics hot path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker (or the timed CPU baseline) -- never as the product path.
 *
 * The oracle follows the reference tree line by line (citations are
 * file:line in the reference): src/dynamics/kernels.cu, integrator.cpp,
 * boundary.cpp (relaxation), src/physics/kessler.cu, src/core/thermo.hpp.
 * It is fed the same padded FP64 base-state / metric / state arrays as the
 * GPU context and keeps the reference's buffer semantics (s_work copy,
 * pi'prev copy, ghost-fill order and staleness).
 *
 * Deliberate, documented deviations from the reference:
 *  - open-boundary zero-gradient ghosts of u (ilast = nx) and v (jlast = ny)
 *    stay inside their row/plane; the reference's m = ng-1 write lands one
 *    element past the row (kernels.cu:757,769) and races with the next row.
 *  - vb: additive base v-wind profile (zero reproduces the reference).
 *  - absmax and the diagnostics use the interior only.
 *
 * Parity unpinned (restated from the reference, matched by the HIP path to
 * 1e-10 in FP64, but checked by no closed form, library routine or
 * reference fixture):
 *  - the values of CUDA's device exp/pow/log/sqrt inside Kessler, Tetens and
 *    the surface physics (host std:: math here, OCML on the GPU);
 *  - in the surface physics (surface.cu:87-387), the scheme's empirical
 *    constants, which nothing outside the reference fixes: the Louis
 *    stability coefficients (9.4, 7.4, 5.3, 4.7), the grey absorption
 *    coefficients (0.22, 0.30 long-wave; 0.008, 0.15 short-wave), the 0.95
 *    clear-sky factor and the lwp / (lwp + 60) cloud albedo, the soil heat
 *    capacity and conductivity fits, the land/sea roughness and albedo, the
 *    local-K mixing length and Ri functions, and the nonlocal boundary-layer
 *    scheme (Ri_c = 0.25, the 6.8 excess, the K profile). Pinned: the
 *    Businger-Dyer functions (defining integrals), the column diffusion and
 *    the soil step (banded library solves, heat conservation), the solar
 *    geometry (declination at the equinoxes and solstices, 15 deg per hour),
 *    the long-wave fluxes (transparent, black and isothermal columns), the
 *    short-wave column (conservation, Beer's law, slant path, solar
 *    constant), the Louis neutral limit and the neutral log-law 10 m / 2 m
 *    diagnostics;
 *  - the seeded ensemble perturbation of the initial state
 *    (main.cpp:105-144), which draws from libstdc++'s
 *    std::normal_distribution<float> (implementation-defined).
 */
#ifndef WFE_ORACLE_H
#define WFE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_model orc_model;

orc_model* orc_create(int nx, int ny, int nz, int ng, double dx, double dy, double dz, double zt);
void orc_destroy(orc_model* m);
/* DynParams fields by name (diff_K, coriolis_use_ub, rayleigh_zd, ...). */
int orc_set_param(orc_model* m, const char* name, double value);
/* Named arrays: base (thb pib rhob dthbdz thbw rhobw thvb thvbw qvb ub vb),
 * metric (zeta_w zeta_c dzeta_c dzeta_w h hx hy hx_u hy_v jac fcor mapf),
 * state s_n (u v w thp pip qv qc qr), tendencies (tu tv tw tthp tpip tqv
 * tqc tqr), rain, boundary (bdy_wgt, bdy_lo_<0..4>, bdy_hi_<0..4>). */
int orc_set_array(orc_model* m, const char* name, const double* data, long n);
int orc_get_array(const orc_model* m, const char* name, double* data, long n);
/* Integrator::init: acoustic coefficients (kernels.cu:915-927). */
int orc_init(orc_model* m);
int orc_set_bdy(orc_model* m, int active, double tfrac);
/* Integrator::step (integrator.cpp:71-111); t is the time at the start of
 * the step (the surface physics uses t + dt). */
int orc_step(orc_model* m, double dt, double t);
/* Integrator::set_physics + SfcPBL::init (surface.cu:440-482): the 2-D
 * arrays tsk, land, lat, lon, soilw (padded, set with orc_set_array first)
 * drive the surface/PBL column physics after every step. */
int orc_set_physics(orc_model* m, int nonlocal, double start_hour_utc, int doy);
int orc_apply_bcs(orc_model* m);
int orc_tendencies(orc_model* m, double dt);
int orc_acoustic(orc_model* m, double dtau, int nsub);
int orc_kessler(orc_model* m, double dt);
double orc_absmax(const orc_model* m, const char* name);
void orc_set_threads(int n);
int orc_get_threads(void);

/* ---- the oracle's own case assembly (case_builder.cpp): Config, file-mode
 * input, Metric, BaseState, initial conditions and BdyManager restated
 * from the reference (src/main.cpp:245-383 and the files it calls). ---- */
typedef struct orc_case orc_case;
/* .ini plus "key=value" overrides; NULL on any configuration/input error. */
orc_case* orc_case_load(const char* ini, int n_over, const char* const* over);
void orc_case_destroy(orc_case* c);
/* nx, ny, nz into dims3; dx, dy, dz into d3. */
int orc_case_dims(const orc_case* c, int* dims3, double* d3);
/* Run scalars (dt, t_end, zt, file_mode) and DynParams by name. */
int orc_case_value(const orc_case* c, const char* name, double* out);
/* Arrays by the product's names (base, metric, initial state, bdy_wgt,
 * bdy_lo_<v>, bdy_hi_<v> for v in u v thp pip qv). */
long orc_case_array_size(const orc_case* c, const char* name);
int orc_case_array(const orc_case* c, const char* name, double* out, long n);
/* BdyManager::update(t) + tfrac(t) (boundary.cpp:83-97) into a model:
 * new targets after a boundary-file crossing, then the time fraction. */
int orc_case_update_bdy(orc_case* c, orc_model* m, double t);

/* Stand-alone pieces for closed-form checks. */
double orc_iface5(double qm2, double qm1, double q0, double q1, double q2, double q3, double vel);
/* Thomas solve of a_k x_{k-1} + b_k x_k + c_k x_{k+1} = d_k, k = 0..n-1,
 * in the elimination order of kernels.cu:589-628. */
void orc_thomas(int n, const double* a, const double* b, const double* c, const double* d,
                double* x);
double orc_qsat_w(double p, double T);
/* Businger-Dyer stability function of zeta = z/L, momentum (heat = 0) or
 * heat (heat = 1) (surface.cu:31-42). */
double orc_sfc_psi(int heat, double zeta);
/* Column-implicit vertical diffusion (surface.cu:61-84) in place of phi;
 * fcg (nz + 1 entries) may be NULL. */
void orc_sfc_diffuse(int nz, double dt, const double* rho, const double* rhow, const double* dzc,
                     const double* dzw, const double* Kw, double src0, double* phi,
                     const double* fcg);
double orc_qsat_i(double p, double T);
/* Surface-physics pieces (surface.cu): cosine of the solar zenith angle;
 * Louis drag/heat coefficients into cd_ch[2]; 10 m wind and 2 m potential
 * temperature into u10_th2[2]; grey long-wave interface fluxes (nz + 1);
 * short-wave layer absorption (returns the surface flux); one implicit step
 * of the four soil layers (Tso[4] in place); soil heat capacity (which = 0)
 * or conductivity (which = 1) at water content w. */
double orc_sfc_cosz(int doy, double utc, double lat, double lon);
void orc_sfc_louis(double Rib, double z1, double z0, double* cd_ch);
void orc_sfc_diag(double V1, double Rib, double z1, double z0, double ths, double th1,
                  double* u10_th2);
void orc_sfc_lw(int nz, const double* T, const double* eps, double Ts, double* Fup, double* Fdn);
double orc_sfc_sw(int nz, double cosz, double lwp, const double* du, const double* dl,
                  double* swabs);
void orc_sfc_soil(double dt, double w, double G, double tdeep, double* Tso);
double orc_sfc_soil_heat(double w, int which);

#ifdef __cplusplus
}
#endif

#endif
