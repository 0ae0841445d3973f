/*
 * sph_amd.h -- C ABI of the MI355X (gfx950) SPH step library libsph_amd.so.
 *
 * Drop-in boundary for the per-step hot path of hassanfv/SPH_3: the Python
 * time loop (sph_code/sph_serial_v2.py:148-218) calls functions of
 * sph_code/libsx.py and sph_code/shear.py; the package sph_3_amd exposes the
 * same Python names and binds these entry points with ctypes.
 *
 * Conventions
 *   - plain C types only; every function returns an int status (SPH_OK = 0);
 *     on failure sph_last_error(ctx) describes it.  Nothing here exits or
 *     aborts the process.
 *   - pos/vel/acc arrays are N x 3 row-major FP64 (numpy C order), scalar
 *     arrays are FP64 of length N, all in the caller's particle order.  The
 *     library never keeps or frees caller memory; results are written into
 *     caller buffers before the call returns.
 *   - one context per host thread; a context owns one HIP stream and is not
 *     re-entrant.  Device data are cached inside the context between calls
 *     (positions, smoothing lengths, masses are compared with the previous
 *     call's), so a time step that calls density, div/curl, acceleration and
 *     du/dt with the same positions sorts, builds the tree and the neighbour
 *     lists once.
 *   - precision: SPH_PREC_MIXED = FP32 pair arithmetic with Kahan-compensated
 *     FP32 sums, FP64 positions/velocities/centres of mass/reductions/h solve;
 *     SPH_PREC_FP64 = the same kernels with FP64 pair arithmetic.
 */
#ifndef SPH_AMD_H
#define SPH_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sph_ctx sph_ctx;

enum sph_status {
    SPH_OK = 0,
    SPH_ERR_ARG = 1,      /* invalid argument (null pointer, n < 1, bad enum) */
    SPH_ERR_HIP = 2,      /* HIP runtime error */
    SPH_ERR_NOCONV = 3,   /* an iterative solve did not converge */
    SPH_ERR_CAPACITY = 4, /* caller buffer too small (required size reported) */
    SPH_ERR_STATE = 5     /* call out of sequence (e.g. sph_step before sph_state_set) */
};

enum sph_precision { SPH_PREC_MIXED = 0, SPH_PREC_FP64 = 1 };
enum sph_softening { SPH_SOFT_SPLINE = 0, SPH_SOFT_PLUMMER = 1 };
enum sph_hmode {
    SPH_H_NEWTON = 0,    /* FP64 Newton-Raphson: (4pi/3)(2h)^3 sum_j W(r_ij,h) = n_ngb */
    SPH_H_REF_COUNT = 1, /* sph_code/libsx.py:804-845 h_smooth_fast: +-0.3 % steps to n_low..n_up */
    SPH_H_KTH = 2        /* sph_code/libsx.py:752-774 do_smoothingX: h = 0.5 * (kth+1)-th distance */
};

/* ---- context ------------------------------------------------------------ */
int sph_ctx_create(int device, sph_ctx **out);
int sph_ctx_destroy(sph_ctx *ctx);
const char *sph_last_error(const sph_ctx *ctx);
int sph_version(void);
/* Device memory held by the context (bytes of every buffer it has allocated: the resident
 * state, the tree, neighbour rows, h-solve scratch, cooling tables, surrogate).  No
 * counterpart in the reference (its arrays are host numpy or fixed CUDA allocations); for
 * capacity planning on one GPU. */
int sph_device_bytes(const sph_ctx *ctx, uint64_t *bytes);
/* Largest size of the smoothing-length solve's candidate distance rows (FP64 r^2 beside each
 * candidate index: 2 KB per target).  A solve whose rows would exceed it keeps the indices
 * only and recomputes the distances from the positions: the same bits, 2 KB per particle
 * less device memory, a slower solve.  The default is a third of the device's memory.  No
 * counterpart in the reference; for runs near the memory capacity of one GPU. */
int sph_set_distance_rows_limit(sph_ctx *ctx, uint64_t bytes);

/* ---- per-function entry points (the reference's Python call surface) ---- */

/* Smoothing lengths.  Replaces sph_code/libsx.py:804 h_smooth_fast(pos, h)
 * (mode SPH_H_REF_COUNT, n_low/n_up = 45/55), libsx.py:754 do_smoothingX
 * ((pos, targets)) (mode SPH_H_KTH, param = 50), and the device h update
 * EV/lib.h:497 smoothing_h_ngb (mode SPH_H_NEWTON, param = n_ngb).
 * targets may be NULL (= pos, m = n); h_in may be NULL for SPH_H_KTH.
 * *n_unconverged counts targets that hit max_steps (mode 1) or did not reach
 * tol (mode 0). */
int sph_smoothing_length(sph_ctx *ctx, int64_t n, const double *pos, int64_t m, const double *targets,
                         const double *h_in, int mode, double param, int n_low, int n_up, int max_steps,
                         double tol, double *h_out, int64_t *n_unconverged);

/* Neighbour lists S_i = { j : |r_i - r_j|^2 < (h_i + h_j)^2 } (self included,
 * FP64, rows ascending).  Replaces EV/ngb.h:125 ngbDB_new_v3.  offsets has
 * n+1 entries and is always written; idx is written when idx_cap >=
 * offsets[n], else SPH_ERR_CAPACITY is returned (idx may be NULL to query). */
int sph_neighbours(sph_ctx *ctx, int64_t n, const double *pos, const double *h, int64_t *offsets,
                   int32_t *idx, int64_t idx_cap);

/* rho_i = sum_j m_j W(r_ij, (h_i+h_j)/2).  Replaces libsx.py:11 getDensity. */
int sph_density(sph_ctx *ctx, int64_t n, const double *pos, const double *m, const double *h, int precision,
                double *rho_out);

/* |div v| and |curl v|.  Replaces shear.py:8 div_curlVel(pos, v, rho, m, h). */
int sph_div_curl(sph_ctx *ctx, int64_t n, const double *pos, const double *vel, const double *rho,
                 const double *m, const double *h, int precision, double *divV_out, double *curlV_out);

/* Pressure + artificial-viscosity acceleration.  divV/curlV NULL: replaces
 * libsx.py:102 getAcc_sph; non-NULL (Balsara switch): replaces shear.py:151
 * getAcc_sph_shear.  acc_out is n x 3. */
int sph_acc_hydro(sph_ctx *ctx, int64_t n, const double *pos, const double *vel, const double *rho,
                  const double *P, const double *c, const double *h, const double *m, const double *divV,
                  const double *curlV, double alpha, int precision, double *acc_out);

/* du/dt.  Replaces libsx.py:355 get_dU (divV/curlV NULL) and shear.py:334
 * get_dU_shear. */
int sph_dudt(sph_ctx *ctx, int64_t n, const double *pos, const double *vel, const double *rho, const double *P,
             const double *c, const double *h, const double *m, const double *divV, const double *curlV,
             double alpha, int precision, double *dudt_out);

/* Self-gravity by tree walk.  Replaces libsx.py:625 getAcc_g_smth(pos, mass,
 * G, epsilon) (soft_kind SPH_SOFT_SPLINE, eps_ij = (eps_i+eps_j)/2, theta = 0
 * gives its exact direct sum) and EV/bh.h:454 compute_forces_kernel.  eps may
 * be NULL (zeros).  n_interactions[0..1] = particle-particle, particle-node
 * (may be NULL).  acc_out is n x 3. */
int sph_gravity(sph_ctx *ctx, int64_t n, const double *pos, const double *m, const double *eps, double G,
                double theta, int soft_kind, double eps_plummer, int precision, double *acc_out,
                int64_t *n_interactions);

/* out5 = [KE, px, py, pz, M] (FP64).  Replaces libsx.py:571 getKE. */
int sph_kinetics(sph_ctx *ctx, int64_t n, const double *vel, const double *m, double *out5);

/* Allocation-free variant of sph_kinetics, same arguments and results.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It stages vel, m and the
 * reduction's partials in a static device array of the library (about 32 MiB per GPU, held
 * until the process ends) instead of context buffers and calls no hipMalloc/hipFree.
 * n <= 1,048,576 (larger n: SPH_ERR_ARG).  Calls that overlap on one GPU share that array
 * and race: serialise them. */
int sph_kinetics_noalloc(sph_ctx *ctx, int64_t n, const double *vel, const double *m, double *out5);

/* Gravitational potential energy PE = -1/2 sum_i m_i phi_i with phi_i = G sum_j m_j psi(r_ij) over
 * the tree walk's interaction list (theta = 0: every pair, exactly the reference's double sum).
 * Replaces libsx.py:585 getPE(pos, m, G, epsilon) (soft_kind SPH_SOFT_SPLINE: its spline potential
 * with eps_ij = (eps_i+eps_j)/2; coincident pairs add nothing, as there); SPH_SOFT_PLUMMER:
 * psi = 1/sqrt(r^2 + eps_plummer^2).  The pair sums run in the precision's type, the total in FP64.
 * phi_out (n values, caller order) may be NULL. */
int sph_potential_energy(sph_ctx *ctx, int64_t n, const double *pos, const double *m, const double *eps, double G,
                         double theta, int soft_kind, double eps_plummer, int precision, double *pe_out,
                         double *phi_out);

/* Host-side check of a particle set, as every entry point that takes particles runs it before
 * uploading anything: finite positions with a non-zero bounding box when n > 1 (coincident
 * particles give no Morton order), and where given finite velocities, masses >= 0, u >= 0,
 * h > 0.  Any array may be NULL.  Returns SPH_OK or SPH_ERR_ARG with the reason in msg
 * (msg_cap bytes, may be 0).  The reference has no such check: its drivers print and exit
 * (EV/ngb.h:254-258, EV/drv.cu:1161-1168) or go on with NaNs.  Needs no GPU. */
int sph_validate_particles(int64_t n, const double *pos, const double *vel, const double *m, const double *u,
                           const double *h, char *msg, int64_t msg_cap);

/* ---- cooling (Particles_in_BH_Tree_III/hcool_with_unit_tests_Here/photolibs3.py and
 *      CR/v4t.h:1768-1884 doCloudyHCooling) ----------------------------------
 * Arrays are elementwise in caller order (cgs units).  J21: amplitude of the
 * J_nu ~ nu^-1 UV field in 1e-21 erg/s/cm^2/Hz/sr (the reference's fixed value
 * is 1e-7); Z: metallicity in solar units (0 = the reference's primordial
 * gas).  A NULL array argument takes the scalar after it. */

/* coolingHeatingRates(T, nHcgs), photolibs3.py:332: heating and cooling per nH^2. */
int sph_cooling_rates(sph_ctx *ctx, int64_t n, const double *T, const double *nH, const double *J21, double J21c,
                      const double *Z, double Zc, double *heat_out, double *lambda_out);
/* what = 0: convert_u_to_temp_h(u, nHcgs, XH), photolibs3.py:379;
 * what = 1: coolingRateFromU_h(u, nHcgs, XH), photolibs3.py:470 (Gamma - Lambda). */
int sph_cooling_from_u(sph_ctx *ctx, int64_t n, int what, const double *u, const double *nH, double XH,
                       const double *J21, double J21c, const double *Z, double Zc, double *out);
/* DoCooling_h(rho, u_old, dt, XH), photolibs3.py:489: implicit update; u_out = -1 where the
 * downward bracket reached u_floor (the reference has no floor). */
int sph_do_cooling(sph_ctx *ctx, int64_t n, const double *rho, const double *u_old, const double *dt, double dtc,
                   double XH, const double *J21, double J21c, const double *Z, double Zc, double u_floor,
                   double *u_out);
/* Build the 4D table of log10 Lambda, log10 Gamma over (log10 nH, log10 T, log10 Z, log10 J21):
 * lo[a], hi[a], npts[a] per axis.  Replaces the absent CR table file HCoolMu.bin (CR/v3t.cu:45). */
int sph_cooling_table_build(sph_ctx *ctx, const double *lo, const double *hi, const int *npts);
/* 4D multilinear lookup (inputs log10, clamped to the table), CR/v4t.h:1681-1884. */
int sph_cooling_table_lookup(sph_ctx *ctx, int64_t n, const double *lognH, const double *logT, const double *logZ,
                             const double *logJ21, double *logL_out, double *logG_out);

/* ---- learned cooling surrogate ----------------------------------------------
 * An MLP 4 -> 128 -> 128 -> 128 -> 1 (ReLU) from the normalised
 * (log10 nH, log10 T, log10 Z, log10 J21) to the normalised log10 Lambda, BF16
 * weights and activations with FP32 accumulation on the gfx950 matrix cores.
 * Its design point is the reference's offline Keras surrogate
 * Particles_in_BH_Tree_III/CLOUDY_replace_CHIMES/HCooling_Here/ML/prog2.py:24-52
 * (4 -> 128 -> 64 -> 1, Adam, MSE); the reference has no device inference path,
 * so these entry points have no one-to-one counterpart there.  The network's
 * box is the cooling table's box; the table (sph_cooling_table_build) is what
 * its output falls back to and what its error map is measured against.
 *
 * Parameters: SPH_MLP_NPARAM floats, flat W1[128][4], b1[128], W2[128][128],
 * b2, W3[128][128], b3, W4[128], b4 (W[out][in]).  norm[10] = box lo[4],
 * hi[4], then mu, sigma of log10 Lambda: x_norm = 2 (v - lo)/(hi - lo) - 1,
 * log10 Lambda = mu + sigma * y_norm. */
#define SPH_MLP_NPARAM 33793

/* Seeded initialisation over the current table's box (mu, sigma from the table). */
int sph_mlp_init(sph_ctx *ctx, uint64_t seed);
/* Load / read parameters and normalisation (loading clears the optimiser state and the error map). */
int sph_mlp_set_params(sph_ctx *ctx, const float *params, const double *norm);
int sph_mlp_get_params(sph_ctx *ctx, float *params, double *norm);
/* Network output y_norm[n] for normalised inputs x_norm[n][4]. */
int sph_mlp_forward(sph_ctx *ctx, int64_t n, const float *x_norm, float *y_norm);
/* Gradient (SPH_MLP_NPARAM floats) of mean_i (y_i - t_i)^2 over the n samples, and that mean. */
int sph_mlp_gradient(sph_ctx *ctx, int64_t n, const float *x_norm, const float *t_norm, float *grad, double *loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected) on the gradient of the n samples. */
int sph_mlp_train_batch(sph_ctx *ctx, int64_t n, const float *x_norm, const float *t_norm, double lr, double *loss);
/* Train on n_samples points drawn uniformly in the box with exact FP64 log10 Lambda targets:
 * `epochs` passes of minibatches of `batch`, Adam with cosine decay from lr to lr/100.
 * loss_out[epochs] (may be NULL) receives the mean normalised squared error of each pass. */
int sph_mlp_train(sph_ctx *ctx, int64_t n_samples, int epochs, int batch, double lr, uint64_t seed, double *loss_out);
/* Error map: the box split into cells[0]*..*cells[3] cells (row-major, last axis fastest),
 * `per` stratified samples each; per cell, in dex: emap_table = max |MLP - table|,
 * emap_exact = max |MLP - exact|, emap_table_exact = max |table - exact| (any may be NULL).
 * Required before sph_mlp_cooling. */
int sph_mlp_calibrate(sph_ctx *ctx, const int *cells, int per, uint64_t seed, float *emap_table, float *emap_exact,
                      float *emap_table_exact);
/* Gated log10 Lambda: the network where the input lies in the box and passes the cell's gate,
 * the table's 4D interpolation elsewhere.  ref = SPH_MLP_REF_EXACT (the Python API's default):
 * fall back where emap_exact > tol [dex] and emap_exact > emap_table_exact (the table measured
 * better there); ref = SPH_MLP_REF_TABLE: fall back where emap_table > tol.  fallback_out[n]
 * (may be NULL) flags the table results; *n_fallback counts them.  Same role as
 * sph_cooling_table_lookup's logL. */
enum sph_mlp_ref { SPH_MLP_REF_EXACT = 0, SPH_MLP_REF_TABLE = 1 };
int sph_mlp_cooling(sph_ctx *ctx, int64_t n, const double *lognH, const double *logT, const double *logZ,
                    const double *logJ21, double tol, int ref, double *logL_out, int32_t *fallback_out,
                    int64_t *n_fallback);
/* Device timing on n resident points uniform in the box, averaged over reps launches:
 * out[0] gated inference ms (ref SPH_MLP_REF_EXACT), out[1] table lookup (log Lambda only) ms, out[2] ungated
 * network ms, out[3] fallback fraction at tol. */
int sph_mlp_timing(sph_ctx *ctx, int64_t n, int reps, uint64_t seed, double tol, double *out);

/* ---- per-step cooling update (CR/v4t.h:1768-1884 doCloudyHCooling, called after the u
 *      update at CR/v3t.cu:1527-1536) ---------------------------------------------
 * In code units: rho_cgs = rho * unit_density_cgs, n_H = XH rho_cgs / m_H, u_cgs = u *
 * unit_u_cgs.  (n_H, T, Z, J21) are clamped into the cooling table's box before the rates
 * are evaluated (doCloudyHCooling clamps to its table bounds); T(u) is the ionisation-
 * equilibrium temperature of photolibs3.py:379-420.  The update is explicit:
 * u_cgs' = max(u_cgs + (Gamma - Lambda) n_H^2 / rho_cgs * dt * unit_time_s, u_floor_cgs),
 * Gamma, Lambda per n_H^2 at the clamped point and n_H unclamped.  Modes:
 * SPH_COOL_EXACT: FP64 T(u) and rates; SPH_COOL_TABLE: T from a temperature table over
 * (log n_H, log u, log J21) and log Lambda, log Gamma from the 4D table;
 * SPH_COOL_MLP: as TABLE with log Lambda from the gated surrogate (sph_mlp_cooling).
 * The table (and for SPH_COOL_MLP the calibrated surrogate) must be loaded in the context.
 * SPH_COOL_NH (resident step only): doCloudyHCooling itself over the table of
 * sph_nh_table_set, with the log N_H of the step's own column-density pass (which must be
 * on) and r = |x|; Z and J21 are not used. */
enum sph_cool_mode { SPH_COOL_EXACT = 0, SPH_COOL_TABLE = 1, SPH_COOL_MLP = 2, SPH_COOL_NH = 3 };
typedef struct sph_cool_params {
    double unit_density_cgs; /* UnitDensity_in_cgs */
    double unit_u_cgs;       /* Unit_u_in_cgs */
    double unit_time_s;      /* unitTime_in_s */
    double XH;               /* hydrogen mass fraction (0.7 in doCloudyHCooling) */
    double u_floor_cgs;      /* 1.8e11 erg/g in doCloudyHCooling */
    double tol;              /* SPH_COOL_MLP: gate tolerance [dex] */
    int mode;                /* sph_cool_mode */
    int ref;                 /* SPH_COOL_MLP: sph_mlp_ref of the gate */
} sph_cool_params;
/* u_out[n] after one step dt (code units) for particles (rho, u, Z, J21) in caller order;
 * fallback_out[n] (may be NULL, SPH_COOL_MLP) flags table results; *n_floor, *n_fallback
 * (may be NULL) count floored particles and table fallbacks.  Replaces doCloudyHCooling. */
int sph_cooling_update(sph_ctx *ctx, int64_t n, const double *rho, const double *u, const double *Z,
                       const double *J21, double dt, const sph_cool_params *cp, double *u_out,
                       int32_t *fallback_out, int64_t *n_floor, int64_t *n_fallback);
/* ---- cooling from a table over (log n_H, log T, r, log N_H): doCloudyHCooling with
 *      find_nx0 and interpolate_4d_hypercube (CR/v4t.h:1450-1476, 1681-1765, 1767-1884; the
 *      table file HCoolMu.bin as CR/v3t.cu:45-122 reads it) ---------------------------------
 * dims = (n_nH, n_T, n_r, n_NH), each >= 2, at most 2048 grid points together; bounds = the
 * (nH, T, r, NH) clamp pairs (low, high); grids ascending; heat, cool (log10 rates), mu and
 * logu (log10 u_cgs) on [nH][T][r][NH], n_nH n_T n_r n_NH floats each.  Copied to the device. */
int sph_nh_table_set(sph_ctx *ctx, const int32_t *dims, const float *bounds, const float *nH_grid,
                     const float *T_grid, const float *r_grid, const float *NH_grid, const float *heat,
                     const float *cool, const float *mu, const float *logu);
/* u_out[n] (code units) after one step dt: particles (pos, rho, u, log_nh = log10 N_H) in caller
 * order, typ (may be NULL) leaves particles with typ != 0 unchanged (the reference's Typ[i] == 0
 * test).  Units, XH and u_floor_cgs come from cp (its mode is not used), gamma is the adiabatic
 * index; *n_floor (may be NULL) counts floored particles.  FP32 arithmetic in the reference's
 * order, with its swapped T and r weights in the interpolation; the reference's float log10
 * and sqrt are taken correctly rounded. */
int sph_nh_cooling(sph_ctx *ctx, int64_t n, const double *pos, const double *rho, const double *u,
                   const double *log_nh, const int32_t *typ, double gamma, double dt, const sph_cool_params *cp,
                   double *u_out, int64_t *n_floor);
/* T(u_cgs, n_H, J21) by the temperature table of SPH_COOL_TABLE / SPH_COOL_MLP (built for XH
 * on first use); the table's counterpart of convert_u_to_temp_h (photolibs3.py:379). */
int sph_cooling_temperature(sph_ctx *ctx, int64_t n, const double *u, const double *nH, const double *J21,
                            double XH, double *T_out);

/* ---- column density from the density gradient --------------------------
 * Replaces grad_rho_norm_ngb (Particles_in_BH_Tree_III/CLOUDY_reloaded_Here/run_simulation/
 * hfvCLibs_v7_hfv_v8_gradRho_v4t.h:716-786) and the host loop after it
 * (…/hfvGpuSphBHtree_gradRho_v3t.cu:1214-1224):
 *   g_i = |sum_{j in S_i} m_j grad W(r_ij, (h_i+h_j)/2)|,
 *   N_H,i = 0.5 (h_i / cbrt(n_ngb) + rho_i / g_i) unit_length_cm * rho_i unit_density_cgs XH / m_H
 * with m_H = 1.673534e-24 g; log_NH = log10 N_H [cm^-2]. */
typedef struct sph_column_params {
    double unit_density_cgs; /* UnitDensity_in_cgs */
    double unit_length_cm;   /* unitLength_in_cm */
    double XH;               /* hydrogen mass fraction (0.7 there) */
    double n_ngb;            /* Nngb_f of h_inter = h / cbrt(Nngb_f) (64 there) */
} sph_column_params;
/* g (may be NULL) and log_NH for particles (pos, m, h, rho) in caller order; the pair sums in
 * the precision's type, the N_H arithmetic in FP64. */
int sph_column_density(sph_ctx *ctx, int64_t n, const double *pos, const double *m, const double *h,
                       const double *rho, const sph_column_params *cp, int precision, double *grad_out,
                       double *log_nh_out);

/* ---- AGN source terms of the outflow driver ------------------------------
 * Isothermal halo (galaxy_isothermal_potential, Particles_in_BH_Tree_III/Evrard_Collapse_2024/
 * hfvCLibs_v7_hfv_v8.1.h:1411-1428): every force evaluation adds a_i += -G M(r) x_i / r^3 with
 * M(r) = 2 sigma^2 r / G, i.e. -2 sigma^2 x_i / r^2, to the total acceleration.
 * Outflow injection (outflow_injector2, same file :1286-1406; after each step,
 * Particles_in_BH_Tree_III/Main_Code/hfvGpuSphBHtree_newCool_v8.cu:1636): the mass M_dot dt plus
 * the leftover buys floor(mass / (2 m_inject)) pairs, the rest carries over; pair j sits at +-r with
 * |r| = r_inject U0 and direction (theta = 2 pi U1, phi = acos(2 U2 - 1)), moves at +-v_in along r,
 * has u = u_inject and the h of the last injected particle (before the first injection: of caller
 * index h_ref_index).  U_k come from a counter-based generator of (seed, step, j, k) instead of
 * cuRAND (see inject_uniform in sph_agn.h).  Injected particles get the next caller indices. */
typedef struct sph_agn_params {
    double sigma;        /* halo velocity dispersion (code units); 0: no halo */
    double M_dot;        /* injected mass per unit time; 0: no injection */
    double v_in;         /* injection speed */
    double m_inject;     /* mass of an injected particle (m_sph_high_res) */
    double u_inject;     /* u of an injected particle (u_for_10K_Temp) */
    double r_inject;     /* injection radius (0.001, one pc, in the driver's units) */
    double Z_inject;     /* metallicity [Z_sun] and J21 of injected particles under cooling */
    double J21_inject;
    uint64_t seed;
    int64_t h_ref_index; /* caller index whose h the first pairs take (ndx_BH - 1 there) */
} sph_agn_params;
/* Room for n_max particles in the resident state, placed by the next sph_state_set (call it
 * first when particles will be injected; an injection beyond the room fails with
 * SPH_ERR_CAPACITY). */
int sph_state_reserve(sph_ctx *ctx, int64_t n_max);
/* Source terms of every later force evaluation (halo) and step (injection); p = NULL: off.
 * Call after sph_state_set (which switches them off). */
int sph_state_set_agn(sph_ctx *ctx, const sph_agn_params *p);
/* Current particle count of the resident state, mass waiting for the next pair, steps since
 * sph_state_set_agn and particles injected since sph_state_set (any pointer may be NULL). */
int sph_state_count(sph_ctx *ctx, int64_t *n, double *leftover_mass, uint64_t *agn_step, int64_t *agn_injected);

/* ---- device-resident time loop (sph_serial_v2.py:148-218 on the GPU) ---- */
typedef struct sph_step_params {
    double dt;          /* fixed step */
    double gamma;       /* adiabatic index */
    double alpha;       /* AV coefficient */
    double G;           /* gravitational constant (0: no gravity) */
    double theta;       /* opening angle */
    double eps_plummer; /* Plummer softening (soft_kind = SPH_SOFT_PLUMMER) */
    double n_ngb;       /* Newton h solve target */
    double h_tol;       /* Newton relative tolerance */
    double period;      /* > 0: side of a periodic box [0, period)^3 with minimum-image distances
                           (positions wrapped by the drift; needs G = 0); 0: open boundaries */
    double C_CFL;       /* > 0: after every force evaluation the per-particle time-step criteria of
                           dt_array_indiv_dt (EV/lib.h:1191-1270) with this Courant factor, the
                           signal velocity's maximum over S_i; the smallest is sph_stats.dt_min and
                           sph_state_get_timestep reads them back.  0: off */
    int soft_kind;      /* SPH_SOFT_SPLINE uses eps_i = h_i (EV/lib.h:584) */
    int precision;      /* SPH_PREC_MIXED / SPH_PREC_FP64 */
    int balsara;        /* 1: shear.py forms, 0: libsx.py forms */
    int timing;         /* 1: record per-phase HIP-event times */
    int energy;         /* 1: the gravity pass also sums the potential, and each force evaluation ends
                           with FP64 totals of the state (sph_stats.energy, momentum, ang_momentum) */
} sph_step_params;

typedef struct sph_stats {
    int64_t n, steps;
    int64_t ngb_total;       /* CSR entries of the last neighbour build */
    int64_t grav_pp, grav_pn; /* interactions of the last gravity pass */
    int64_t h_unconverged;   /* cumulative */
    int64_t h_rounds;        /* rounds of the last h solve */
    double phase_ms[20];     /* accumulated HIP-event milliseconds per phase (sph_phase) */
    int64_t phase_calls[20];
    int64_t h_tests;         /* last h solve: FP64 distance tests of its walks */
    int64_t h_terms;         /* last h solve: kernel terms summed by its Newton iterations */
    int64_t ngb_extra;       /* last step: neighbour pairs found only from the other particle's candidates */
    int64_t cool_floor;      /* last step: particles the cooling update set to u_floor */
    int64_t cool_fallback;   /* last step: table fallbacks of the surrogate (SPH_COOL_MLP) */
    int64_t h_big;           /* last h solve: targets whose solution needs more candidates than a
                                row holds (solved by a candidate-free tree-walk Newton) */
    double dt_min;           /* last force evaluation with C_CFL > 0: min_i dt_i (else 0) */
    /* last force evaluation with sph_step_params.energy (FP64 sums at the end of the evaluation:
     * positions of the forces, velocities after the second kick, u after its update and cooling;
     * the priming evaluation: the initial velocities).  energy = {KE = sum m v^2/2, U = sum m u,
     * PE = -1/2 sum m phi of the self-gravity tree walk (0 without gravity; the AGN halo's external
     * potential is not included)}; momentum = sum m v; ang_momentum = sum m x cross v. */
    double energy[3];
    double momentum[3];
    double ang_momentum[3];
} sph_stats;

/* phase indices of sph_stats.phase_ms (SPH_PH_HMOMENTS: the node boxes built for the h walk;
 * SPH_PH_DIVCURL stays empty in the Balsara form, whose div/curl v share the density pass) */
enum sph_phase {
    SPH_PH_KICKDRIFT = 0, SPH_PH_ORDER, SPH_PH_TOPOLOGY, SPH_PH_HSOLVE, SPH_PH_MOMENTS, SPH_PH_NGB,
    SPH_PH_DENSITY, SPH_PH_GRAVITY, SPH_PH_DIVCURL, SPH_PH_ACCSPH, SPH_PH_KICK, SPH_PH_DUDT,
    SPH_PH_UUPDATE, SPH_PH_HMOMENTS, SPH_PH_COOLING, SPH_PH_COLUMN, SPH_PH_TIMESTEP, SPH_PH_COUNT
};

/* Load the resident state (caller order); u and h are per particle. */
int sph_state_set(sph_ctx *ctx, int64_t n, const double *pos, const double *vel, const double *m,
                  const double *u, const double *h);
/* Read it back in caller order; any pointer may be NULL. */
int sph_state_get(sph_ctx *ctx, double *pos, double *vel, double *u, double *h, double *rho, double *acc);
/* Per-particle metallicity Z [Z_sun] and UV amplitude J21 (caller order) of the resident
 * state and the cooling of every later sph_step, applied after its u update
 * (CR/v3t.cu:1527-1536); cp = NULL switches cooling off.  Call after sph_state_set. */
int sph_state_set_cooling(sph_ctx *ctx, const double *Z, const double *J21, const sph_cool_params *cp);
/* Column density of every later force evaluation (after its density pass, as the reference
 * computes it each step, hfvGpuSphBHtree_gradRho_v3t.cu:1204-1226); cp = NULL switches it
 * off.  Call after sph_state_set (which switches it off).  sph_state_get_column reads the last
 * one back in caller order (either pointer may be NULL). */
int sph_state_set_column(sph_ctx *ctx, const sph_column_params *cp);
int sph_state_get_column(sph_ctx *ctx, double *grad_out, double *log_nh_out);
/* Per-particle dt_i and dh/dt_i = h div v / 3 of the last force evaluation with C_CFL > 0
 * (caller order; particles injected after it read NaN); either pointer may be NULL. */
int sph_state_get_timestep(sph_ctx *ctx, double *dt_out, double *dhdt_out);
/* Restart: sph_state_get_duprev reads the du/dt the next u update averages with (caller order);
 * sph_state_set_restart, after sph_state_set (and sph_state_set_agn / _cooling when used), loads
 * the last force evaluation's acceleration and that du/dt (caller order), the injection's
 * leftover mass, step counter and injected count, and marks the state primed, so sph_step
 * continues the saved run.  The reference has no resume; its snapshots lack these fields. */
int sph_state_get_duprev(sph_ctx *ctx, double *duprev_out);
int sph_state_set_restart(sph_ctx *ctx, const double *acc, const double *duprev, double agn_leftover,
                          uint64_t agn_step, int64_t agn_injected);
/* Forces at the current positions and the first u update (sph_serial_v2.py:94-143). */
int sph_prime(sph_ctx *ctx, const sph_step_params *p);
/* One KDK step (sph_serial_v2.py:148-203), entirely on the device. */
int sph_step(sph_ctx *ctx, const sph_step_params *p);
/* Neighbour sets S_i of the resident state at its positions and h (last sph_prime or
 * sph_step), in the caller's numbering and the layout of sph_neighbours (rows ascending);
 * the step builds them from its smoothing-length candidates, this reads them back. */
int sph_state_neighbours(sph_ctx *ctx, int64_t *offsets, int32_t *idx, int64_t idx_cap);
int sph_get_stats(sph_ctx *ctx, sph_stats *out);
int sph_reset_stats(sph_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* SPH_AMD_H */
