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
int orc_step(orc_model* m, double dt, double t);
int orc_apply_bcs(orc_model* m);
int orc_tendencies(orc_model* m, double dt);
int orc_acoustic(orc_model* m, double dtau, int nsub);
int orc_kessler(orc_model* m, double dt);
double orc_absmax(const orc_model* m, const char* name);
void orc_set_threads(int n);
int orc_get_threads(void);

/* Stand-alone pieces for closed-form checks. */
double orc_iface5(double qm2, double qm1, double q0, double q1, double q2, double q3, double vel);
/* Thomas solve of a_k x_{k-1} + b_k x_k + c_k x_{k+1} = d_k, k = 0..n-1,
 * in the elimination order of kernels.cu:589-628. */
void orc_thomas(int n, const double* a, const double* b, const double* c, const double* d,
                double* x);
double orc_qsat_w(double p, double T);
double orc_qsat_i(double p, double T);

#ifdef __cplusplus
}
#endif

#endif
