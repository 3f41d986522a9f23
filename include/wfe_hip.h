/*
 * wfe_hip.h -- C ABI of the MI355X (gfx950) dynamics hot path.
 *
 * Plain C: POD structs, host pointers and sizes, int status codes. Every
 * function returns WFE_OK (0) on success or a negative WFE_E* code; the text
 * of the last error on the calling thread is in wfe_last_error().
 *
 * All 3-D host arrays use the reference's padded layout
 *   idx(i,j,k) = ((k+ng)*NY + (j+ng))*NX + (i+ng),
 *   NX = nx+2ng, NY = ny+2ng, NZ = nz+1+2ng   (src/core/grid.hpp:22-36),
 * as float64 on the host whatever the device precision. 2-D arrays are
 * NX*NY (padded), 1-D vertical arrays NZ (index k+ng).
 *
 * The reference has no foreign-function interface; each entry point below
 * names the C++ interface of the reference it replaces (file:line in the
 * reference tree).
 */
#ifndef WFE_HIP_H
#define WFE_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WFE_HIP_ABI_VERSION 1

enum {
  WFE_OK = 0,
  WFE_E_ARG = -1,     /* invalid argument, size or state */
  WFE_E_HIP = -2,     /* HIP runtime error */
  WFE_E_CONFIG = -3,  /* case configuration error */
  WFE_E_IO = -4,      /* file read/write error */
  WFE_E_NOMEM = -5    /* allocation failure */
};

/* Device precision of a context. */
enum {
  WFE_PREC_FP64 = 0,  /* everything stored and computed in FP64 ("plain FP64") */
  WFE_PREC_MIXED = 1  /* FP32 storage and bulk arithmetic; pi' stored in FP64;
                         acoustic column assembly and tridiagonal elimination,
                         pi = pibar + pi' and the pi' gradient in FP64 */
};

/* Field selectors for upload/download/absmax. */
enum {
  WFE_F_U = 0, WFE_F_V, WFE_F_W, WFE_F_THP, WFE_F_PIP, WFE_F_QV, WFE_F_QC, WFE_F_QR,
  /* Slow tendencies. WFE_F_TPIP is the Davies rim relaxation term of pi'
   * (zero elsewhere, kernels.cu:901,981-995): the substeps read it only in
   * columns with a positive rim weight. */
  WFE_F_TU = 8, WFE_F_TV, WFE_F_TW, WFE_F_TTHP, WFE_F_TPIP, WFE_F_TQV, WFE_F_TQC, WFE_F_TQR,
  WFE_F_RAIN = 16, /* 2-D accumulated surface rain [kg m-2] */
  /* 2-D surface-physics fields (after wfe_set_physics): skin temperature
   * [K], boundary-layer height [m], 2 m temperature [K], 10 m wind [m/s] */
  WFE_F_TSK = 17, WFE_F_PBLH, WFE_F_T2M, WFE_F_U10,
  WFE_F_COUNT = 21
};

/* GDims (src/core/grid.hpp:22-36). ng must be 3. */
typedef struct {
  int nx, ny, nz, ng;
  double dx, dy, dz;
} wfe_grid;

/* DynParams (src/dynamics/params.hpp:8-25). Booleans are 0/1. */
typedef struct {
  double diff_K;
  double coriolis_f;
  int coriolis_use_ub;
  double rayleigh_zd;
  double rayleigh_alpha;
  int acoustic_ns;
  double acoustic_beta;
  double acoustic_smdiv;
  int bc_x_open;
  int bc_y_open;
  double cstar;
  int moisture;
  int microphysics;
  int pd_moist;
  int w_damping;
  double w_abort;
} wfe_params;

/* DevProf (src/core/base_state.hpp:16-27): host FP64 arrays. 3-D fields are
 * padded (NX*NY*NZ); ub and vb are 1-D [NZ]. vb is an additive extension of
 * the reference (base v wind); NULL means zero. */
typedef struct {
  const double *thb, *pib, *rhob, *dthbdz, *thbw, *rhobw, *thvb, *thvbw, *qvb;
  const double *ub, *vb;
} wfe_base;

/* DevMetric (src/core/metric.hpp:16-30): host FP64 arrays. */
typedef struct {
  const double *zeta_w, *zeta_c, *dzeta_c, *dzeta_w;             /* [NZ]    */
  const double *h, *hx, *hy, *hx_u, *hy_v, *jac, *fcor, *mapf;   /* [NX*NY] */
  double zt;
} wfe_metric;

typedef struct wfe_ctx wfe_ctx;

/* ---- integrator context: Integrator (src/dynamics/integrator.hpp:18-42) ---- */

/* Integrator::init (integrator.cpp:43-69): allocates device state, uploads
 * base state and metric, precomputes the acoustic coefficients
 * (kernels.cu:915-927). The state starts at zero. */
int wfe_create(const wfe_grid* g, const wfe_params* p, const wfe_base* base,
               const wfe_metric* m, int precision, wfe_ctx** out);
int wfe_destroy(wfe_ctx* ctx);

/* Integrator::step (integrator.cpp:71-111): one WS2002 RK3 large step with
 * split-explicit acoustic substeps, moisture stage updates, boundary
 * relaxation (when set), Kessler microphysics (when moisture and
 * microphysics are on) and the surface physics (when set). t is the
 * simulation time at the start of the step: the surface physics takes its
 * hour of day at t + dt; the boundary time fraction is set separately with
 * wfe_set_bdy_tfrac. Asynchronous with respect to the host. */
int wfe_step(wfe_ctx* ctx, double dt, double t);

/* SfcPBL::init (src/physics/surface.cu:440-482) + Integrator::set_physics
 * (src/dynamics/integrator.hpp:27): surface/PBL/soil/radiation column
 * physics (surface.cu:87-437) applied to the state after every step.
 * tsk, land, lat, lon, soilw: FP64 host arrays on the padded 2-D plane
 * (NX*NY, n2 elements; interior used): skin temperature [K], land mask (> 0.5
 * land), latitude and longitude [deg], soil moisture [m3 m-3]. Deep soil and
 * the four soil layers start at tsk. start_hour_utc: hour of day of t = 0;
 * doy: day of year (1..366); nonlocal: 1 for the nonlocal (Troen-Mahrt /
 * Hong-Pan) boundary layer, 0 for local-Ri diffusivity. The skin
 * temperature, boundary-layer height, 2 m temperature and 10 m wind are
 * read back with WFE_F_TSK .. WFE_F_U10. Ensemble steps do not advance the
 * hour of day of members with surface physics. */
typedef struct {
  const double *tsk, *land, *lat, *lon, *soilw;
  double start_hour_utc;
  int doy;
  int nonlocal;
} wfe_sfc;
int wfe_set_physics(wfe_ctx* ctx, const wfe_sfc* sfc, size_t n2);

/* ---- tangent-linear and adjoint of the large step ----
 * No counterpart in the reference. Scope: FP64 or mixed contexts without
 * microphysics or surface physics (WFE_E_CONFIG otherwise); terrain, map
 * factors, w damping, PD q_v transport (limiter branches frozen at the
 * trajectory), open boundaries (radiation, branch
 * frozen at the trajectory), Davies relaxation (targets and time fraction
 * held at the context's current values over a gradient's steps), Coriolis
 * and constant-K diffusion are included; with moisture on, q_v is linearised
 * (the _q entries); nz <= 63, acoustic_ns <= 16. In a mixed context the
 * linearisation is about the FP32 trajectory the context computes (FP32
 * production); perturbations, adjoint variables and J are FP64 in both.
 * Fields are (u, v, w, theta', pi') as five FP64 host
 * arrays in the padded reference layout (n = padded 3-D element count);
 * ghost values of inputs are ignored (they are filled from the interior).
 * The upwind branch of every interface value is taken from the trajectory.
 *
 * wfe_tl_step: dy = M'(x) dx for one large step M around the context's
 *   current state x, which is left unchanged.
 * wfe_ad_step: xb = M'(x)^T yb, same conventions.
 * wfe_adj_gradient: J = (1/N) sum_{n=1..N} sum_{i,j} wgt(i,j) (u^2 + v^2)/2
 *   at the lowest level over N steps from the current state (wgt: padded
 *   2-D array, n2 elements), and dJ/d(initial u, v, w, theta', pi') in grad;
 *   the state is kept every ckpt_every steps and each segment recomputed.
 *   The context's state is left unchanged. */
int wfe_tl_step(wfe_ctx* ctx, double dt, const double* const* dx, double* const* dy, size_t n);
int wfe_ad_step(wfe_ctx* ctx, double dt, const double* const* yb, double* const* xb, size_t n);
int wfe_adj_gradient(wfe_ctx* ctx, double dt, int nsteps, int ckpt_every, const double* wgt2d,
                     size_t n2, double* J, double* const* grad, size_t n);
/* The same with q_v as a sixth field (u, v, w, theta', pi', q_v): with
 * moisture on (microphysics off, q_c = q_r = 0) q_v is transported (flux
 * form, or positive-definite with the limiter's branches frozen), adds
 * g 0.61 q_v' to the w buoyancy and scales theta_v by (1 + 0.61 q_v), and is
 * linearised with the rest; without moisture the q_v input is ignored and
 * its output is zero. The five-field entries above are these with q_v
 * perturbations zero and the q_v output dropped. No counterpart in the
 * reference. */
int wfe_tl_step_q(wfe_ctx* ctx, double dt, const double* const* dx, double* const* dy,
                  size_t n);
int wfe_ad_step_q(wfe_ctx* ctx, double dt, const double* const* yb, double* const* xb,
                  size_t n);
int wfe_adj_gradient_q(wfe_ctx* ctx, double dt, int nsteps, int ckpt_every,
                       const double* wgt2d, size_t n2, double* J, double* const* grad,
                       size_t n);

/* ---- ensembles: members of any grid size advancing together ----
 * No counterpart in the reference (which runs one case per process). An
 * ensemble borrows n existing contexts (each from wfe_create or
 * wfe_case_create_ctx; all must be at the same step parity, i.e. created
 * together or stepped the same number of times) and advances all of them by
 * one large step per wfe_ensemble_step, member q with dt[q]: one hipGraph
 * holds every member's step as an independent branch, one launch per step.
 * Batched substeps (one launch of each acoustic substep kernel over every
 * member's grid) are switched off; WFE_ENSEMBLE_BATCH=1 makes
 * wfe_ensemble_create fail with WFE_E_ARG. Members with surface physics
 * (wfe_set_physics) need the time of each step: wfe_ensemble_step_at takes
 * t[q], the time at the start of member q's step (as wfe_step's t), and
 * wfe_ensemble_step fails with WFE_E_ARG for such an ensemble.
 * The first call runs each member's first (tuning) step eagerly. Work issued
 * afterwards on a member context (download, absmax, its own wfe_step) is
 * ordered after the ensemble step. Destroy the ensemble before its members. */
typedef struct wfe_ensemble wfe_ensemble;
int wfe_ensemble_create(wfe_ctx* const* members, int n, wfe_ensemble** out);
int wfe_ensemble_destroy(wfe_ensemble* e);
int wfe_ensemble_step(wfe_ensemble* e, const double* dt);
int wfe_ensemble_step_at(wfe_ensemble* e, const double* dt, const double* t);
int wfe_ensemble_sync(wfe_ensemble* e);
/* How many members the last captured ensemble step advanced in batched
 * substep launches (0 while batching is switched off). */
int wfe_ensemble_batched(const wfe_ensemble* e, int* members);

/* Field3D::upload / download (src/core/field3d.hpp:42-47) of the time-n
 * state (Integrator::state(), integrator.hpp:24), of the slow tendencies, of
 * the 2-D rain (Integrator::rain(), integrator.hpp:25) or (download only) of
 * the 2-D surface fields (SfcPBL::tsk/pblh/t2m/u10, surface.hpp:26-29). n is
 * the element count (npts, or NX*NY for 2-D fields). Synchronous. */
int wfe_upload(wfe_ctx* ctx, int field, const double* host, size_t n);
int wfe_download(wfe_ctx* ctx, int field, double* host, size_t n);

/* field_absmax (kernels.hpp:69-71, kernels.cu:967-979): NaN-safe max|f| over
 * the interior (w: levels 0..nz). NaN/Inf report +Inf. Synchronous. */
int wfe_absmax(wfe_ctx* ctx, int field, float* out);

/* ---- output path ---- */

/* The driver's diagnostics record (Diag, src/main.cpp:197-200). */
typedef struct {
  double wmax;    /* max |w| over the interior, levels 0..nz */
  double thmin;   /* min(0, theta'), levels 0..nz-1 */
  double thmax;   /* max(0, theta') */
  double pipmax;  /* max |pi'| */
  double qcmax;   /* max(0, qc); 0 for dry contexts */
  double qrmax;   /* max(0, qr); 0 for dry contexts */
  int finite;     /* 0 when any w, theta', pi' or qr value is NaN or Inf */
} wfe_diag;

/* diagnose (src/main.cpp:202-240) as a device reduction over the interior
 * of the time-n state: the same extremes, start values and NaN rules as the
 * host loops (NaN never enters an extreme, it clears `finite`; Inf enters),
 * with only the record copied back instead of five whole fields.
 * Synchronous. */
int wfe_diagnose(wfe_ctx* ctx, wfe_diag* out);

/* diagnose (src/main.cpp:202-240) over caller-owned device fields, as the
 * reference's diagnose takes any State: w, theta', qc, qr are float (MIXED)
 * or double (FP64) device arrays and pi' a double device array, all in the
 * reference's padded layout (idx above: row length nx+2ng, ng = 3); qc and
 * qr both NULL for a dry state. stream is a hipStream_t (NULL: the null
 * stream) ordered after the writes of the fields. Allocates its workgroup
 * partial records on the device for the call. Synchronous. */
int wfe_diagnose_fields(const wfe_grid* g, int precision, const void* w, const void* thp,
                        const double* pip, const void* qc, const void* qr, void* stream,
                        wfe_diag* out);

/* wfe_diagnose_fields without any allocation. No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, hosts that forbid hipMalloc).
 * The workgroup records go to a fixed 1 MiB device workspace inside the
 * library, one per GPU, held until the process ends. Limit: at most 16384
 * workgroups of 64 x 4 columns x 8 levels, i.e. ceil(nx/64) * ceil(ny/4) *
 * ceil((nz+1)/8) <= 16384; larger grids are rejected with WFE_E_ARG. Calls
 * on the same GPU share the workspace: calls that overlap in time on one
 * GPU (several host threads or streams) race and must be serialised by the
 * caller. Same arguments, result and synchronisation as
 * wfe_diagnose_fields. */
int wfe_diagnose_fields_noalloc(const wfe_grid* g, int precision, const void* w,
                                const void* thp, const double* pip, const void* qc,
                                const void* qr, void* stream, wfe_diag* out);

/* Writer::write_field (src/io/writer.cpp:89-107) without the full download:
 * packs the interior of a state field (WFE_F_U .. WFE_F_QR) or the rain
 * (WFE_F_RAIN) as float32 in the output order (i fastest, k slowest; nz
 * levels, nz+1 for w, 1 for the rain; n must be nx*ny*levels) on the
 * context's stream, then copies it into dst on a copy stream. Returns at
 * once; dst is written once the copy has run (see wfe_output_wait). Later
 * steps do not wait for the copy. dst from wfe_host_alloc (pinned) makes the
 * copy asynchronous; other host memory works but the copy then blocks. */
int wfe_pack_field_async(wfe_ctx* ctx, int field, float* dst, size_t n);
/* Ticket for every copy enqueued so far by wfe_pack_field_async. */
int wfe_output_mark(wfe_ctx* ctx, long long* ticket);
/* Blocks until the copies of a ticket are in host memory. Callable from a
 * thread other than the one stepping the context. The last 64 tickets are
 * valid. */
int wfe_output_wait(wfe_ctx* ctx, long long ticket);
/* Device buffers for callers without a HIP runtime of their own (the
 * Python tools load this library's runtime; a second HIP runtime in the
 * same process cannot share the GPU): allocation, release and synchronous
 * copies to and from host memory. */
int wfe_device_alloc(size_t bytes, void** out);
int wfe_device_free(void* p);
int wfe_device_upload(void* dev, const void* host, size_t bytes);
int wfe_device_download(void* host, const void* dev, size_t bytes);
/* Pinned (page-locked) host memory for wfe_pack_field_async destinations. */
int wfe_host_alloc(size_t bytes, void** out);
int wfe_host_free(void* p);

/* cudaDeviceSynchronize as used by the driver (main.cpp:441,468). */
int wfe_sync(wfe_ctx* ctx);

/* ---- boundary relaxation: BdyManager (src/dynamics/boundary.hpp:16-38) ---- */

/* BdyManager::init weights (boundary.cpp:36-54): 2-D relaxation coefficient
 * [1/s] on the padded NX*NY grid; activates relaxation in wfe_step. */
int wfe_set_bdy_weights(wfe_ctx* ctx, const double* wgt, size_t n);
/* BdyManager::load_into (boundary.cpp:51-81): uploads relaxation target
 * var (0 u, 1 v, 2 theta', 3 pi', 4 qv) into slot 0 (lo) or 1 (hi). */
int wfe_set_bdy_target(wfe_ctx* ctx, int slot, int var, const double* host, size_t n);
/* BdyManager::update's lo <- hi swap (boundary.cpp:83-91). */
int wfe_swap_bdy_targets(wfe_ctx* ctx);
/* BdyManager::tfrac (boundary.cpp:93-97) for the next steps. */
int wfe_set_bdy_tfrac(wfe_ctx* ctx, double tfrac);

/* ---- phase-level entry points: free functions of kernels.hpp:14-71 ---- */

/* apply_bcs (kernels.cu:805-835) on the time-n state. */
int wfe_apply_bcs(wfe_ctx* ctx);
/* compute_mass_fluxes + compute_divergence + compute_tendencies
 * (kernels.cu:837-913) from the time-n state into the tendency fields,
 * after apply_bcs; dt is the large step (PD limiter, w damping). */
int wfe_compute_tendencies(wfe_ctx* ctx, double dt);
/* nsub successive acoustic_substep calls (kernels.cu:929-965) started from
 * the time-n state with the current tendencies, exactly as one RK3 stage
 * runs them (fast system restarted from time n, pi'prev = pi'(n));
 * the result replaces the time-n state. */
int wfe_acoustic_substeps(wfe_ctx* ctx, double dtau, int nsub);
/* kessler_step (src/physics/kessler.hpp:17-18) on the time-n state. */
int wfe_kessler(wfe_ctx* ctx, double dt);

/* ---- measurement (the reference's WFE_PROF section timer,
 *      src/dynamics/integrator.cpp:18-41, at kernel granularity) ---- */

/* mode 1: record a HIP event pair on the context's stream around every
 * kernel launch; 0: stop; 2: clear the totals and start. */
int wfe_kernel_timing(wfe_ctx* ctx, int mode);
/* Per-kernel totals as text lines "name launches total_ms\n" (synchronous). */
int wfe_kernel_timing_report(wfe_ctx* ctx, char* buf, size_t n);
/* Average launch time [us] of one kernel of the step ("acou_wpi", "acou_uv",
 * "acoustic_fused", "tendencies", "massflux", "fill_ghosts", "stage_q"),
 * launched reps times back to back on the context's stream between two HIP
 * events, with the arguments it has in the first substep / stage of a step.
 * Only work buffers, tendencies, fluxes and ghost cells are written, all of
 * which the next step recomputes. A measurement helper beside the profile
 * above; the reference has no counterpart. */
int wfe_kernel_bench(wfe_ctx* ctx, const char* kernel, int reps, double* avg_us);

/* Device-wide streaming copy of `bytes` (rounded down to 16 B), repeated
 * reps times on a private stream: achieved read+write bandwidth in GB/s.
 * No counterpart in the reference; used to state the measured HBM peak. */
int wfe_copy_probe(size_t bytes, int reps, double* gbps);

/* ---- introspection ---- */
const char* wfe_last_error(void);
int wfe_abi_version(void);
/* Device name of the context's GPU into buf (for run_info.txt, main.cpp:404). */
int wfe_device_name(char* buf, size_t n);

/* ---- host case builder: Config + Metric + BaseState + initial conditions
 *      exactly as src/main.cpp:248-383 assembles them ---- */

typedef struct wfe_case wfe_case;

/* Loads an .ini case (Config::load, config.cpp:17-31) with "key=value"
 * overrides (main.cpp:254-259), builds the grid, DynParams, Metric,
 * BaseState, file-mode input and the initial state. */
int wfe_case_load(const char* ini_path, int n_over, const char* const* over, wfe_case** out);
int wfe_case_destroy(wfe_case* c);
int wfe_case_grid(const wfe_case* c, wfe_grid* g);
int wfe_case_params(const wfe_case* c, wfe_params* p);
/* Run keys: dt, t_end, out_interval, diag_interval, zt; surface physics:
 * sfc_on, sfc_start_hour, sfc_doy, sfc_nonlocal. */
int wfe_case_scalar(const wfe_case* c, const char* name, double* out);
/* Element count of a named host array, or 0 if unknown. Names: base fields
 * (thb pib rhob dthbdz thbw rhobw thvb thvbw qvb ub vb), metric fields
 * (zeta_w zeta_c dzeta_c dzeta_w h hx hy hx_u hy_v jac fcor mapf), initial
 * state (u v w thp pip qv qc qr), boundary (bdy_wgt, bdy_lo_<v>, bdy_hi_<v>
 * for v in u v thp pip qv), surface physics (sfc_tsk sfc_land sfc_lat
 * sfc_lon sfc_soilw). */
size_t wfe_case_array_size(const wfe_case* c, const char* name);
int wfe_case_array(const wfe_case* c, const char* name, double* out, size_t n);
/* Integrator::init + IC upload (+ BdyManager::init in file mode, + the
 * surface physics with physics = simple). */
int wfe_case_create_ctx(const wfe_case* c, int precision, wfe_ctx** out);
/* BdyManager::update(t) then tfrac(t) for the context (boundary.cpp:83-97),
 * reading further boundary files when t passes the next boundary time. Each
 * context created from the case keeps its own bracket of boundary times, so
 * one case can drive several contexts that stand at different times. */
int wfe_case_update_bdy(wfe_case* c, wfe_ctx* ctx, double t);
/* Unused-key list (Config::unused, config.cpp:51-60), newline-separated. */
int wfe_case_unused_keys(const wfe_case* c, char* buf, size_t n);

#ifdef __cplusplus
}
#endif

#endif /* WFE_HIP_H */
