"""Throughput benchmark of the SPH step on one MI355X.

Workload: Evrard collapse, N = 1,048,576 equal-mass gas particles (M = 1,
R = 1, G = 1, rho ~ 1/r from r = R sqrt(U), seeded isotropic directions,
u = 0.05, gamma = 5/3), cubic spline, N_ngb = 64 (FP64 Newton h solve),
theta = 0.5, Plummer eps = 1e-3, alpha = 1 with the Balsara switch,
dt = 1e-3.  One "step" is the full KDK step of sph_serial_v2.py on the
device: kick/drift, Morton sort, tree, h solve, neighbour rows,
density+EOS+div/curl, tree gravity, acceleration, kick, du/dt, u update.

Prints ONE JSON line: particle-steps/s (whole job, inputs resident in HBM),
the roofline of the dominant kernel from HIP events on the library's stream
(HBM traffic per launch from the committed PMC summary), the all-FP64 form
of the same kernels, the CPU baseline (the oracle's FP64 all-pairs
restatement of the reference step, row-sampled) and the cooling surrogate at
4,194,304 particles (training, gated inference, MFMA roofline, fallback rate).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
FLOP_PER_INTERACTION = 20
# h solve (FP64): a walk distance test is 3 sub + 3 mul + 2 add; a Newton kernel
# term is q = r/h (1), (2-q)+, (1-q)+ (4), squares (2), w (1 mul + 2 fma) and
# q w'(q) (1 mul + 2 fma) = 16 flop (fma = 2)
H_FLOP_PER_TEST = 8
H_FLOP_PER_TERM = 16


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def per_particle_bytes(kbar):
    """Algorithmic (compulsory) HBM bytes per particle of each pass with this
    build's layout: int32 CSR index stream (4 B per neighbour), int64 row
    offsets, FP64 x,y,z / v / m / h, FP32 hydro scalars (mixed form), 128-B
    tree node records (2N-1 of them)."""
    idx = 4.0 * kbar + 8.0
    return {
        "order": 6 * (8 + 4) * 2 + 24 + 13 * 8 * 2,          # 6 radix passes of (key,val) r+w, keys, 13 arrays gathered
        "moments": 24 + 8 + 8 + 2 * 128,                     # x, m, h in; ~2 node records out
        "hsolve": 24 + 8 + 8 + 2 * 128,                      # x, h in, h out, tree read once
        "ngb": 24 + 8 + 2 * 128 + idx,                       # x, h, tree in; CSR out
        # density fused with div/curl v (Balsara form): x,v,m,h,u in; rho,P,c,P/rho^2,divV,curlV,f out
        "density": idx + 24 + 24 + 8 + 8 + 8 + 4 * 4 + 3 * 4,
        "divcurl": idx + 24 + 24 + 8 + 8 + 4 + 4 + 3 * 4,    # separate pass (libsx form): x,v,m,h,rho,c in
        "accsph": idx + 24 + 24 + 8 + 8 + 4 * 4 + 24 + 24,   # x,v,m,h,rho,P/rho^2,c,f,a_grav in; a out
        "dudt": idx + 24 + 24 + 8 + 8 + 4 * 4 + 4,           # ... in; du out
        "kick_drift": 6 * 8 * 2 + 3 * 8 + 3 * 8,
        "kick": 3 * 8 * 2 + 3 * 8,
        "u_update": 8 * 2 + 4 + 8 * 2,
    }


KERNEL_OF_PHASE = {"order": "k_rs_scatter+k_gather", "moments": "k_moments_local+k_moments_up",
                   "hsolve": "k_hsolve", "ngb": "k_ngb<2>", "density": "k_density_divcurl<float,16>",
                   "divcurl": "k_divcurl<float,16>", "accsph": "k_accsph<float,16>", "dudt": "k_dudt<float,16>",
                   "gravity": "k_gravity_group<float,2>", "moments_h": "k_moments_local+k_moments_up"}


def cpu_baseline(sim_state, n, m, rows, seed=0):
    """Time the oracle's FP64 all-pairs restatement of the reference step
    (libsx.h_smooth_fast, getDensity, getAcc_g_smth, shear.div_curlVel,
    getAcc_sph_shear, get_dU_shear) on `rows` sampled target rows against all
    n sources; extrapolate linearly in rows (rows are independent)."""
    import oracle as O

    O.lib()
    rng = np.random.default_rng(seed)
    pos = sim_state["pos"]; vel = sim_state["vel"]; h = sim_state["h"]; u = sim_state["u"]
    rho = sim_state["rho"]
    P = (2.0 / 3.0) * rho * u
    c = np.sqrt(10.0 / 9.0 * u)
    eps = np.full(n, 1e-3)
    # sampled rows are contiguous after a random shuffle of the particle order,
    # which keeps the reference's row-range interface
    perm = rng.permutation(n)
    pos_p, vel_p, h_p, m_p = pos[perm], vel[perm], h[perm], m[perm]
    rho_p, P_p, c_p = rho[perm], P[perm], c[perm]
    dv = np.zeros(n); cv = np.zeros(n)
    t = {}
    t0 = time.perf_counter()
    O.h_smooth_fast(pos_p, h_p * (50.0 / 64.0) ** (1 / 3), nbeg=0, nend=rows)
    t["h_smooth_fast"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.density_allpairs(pos_p, m_p, h_p, 0, rows)
    t["getDensity"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.grav_direct(pos_p, m_p, eps, 1.0, O.SOFT_SPLINE, 0.0, 0, rows)
    # the reference's symmetric j > i loop does half the pairs per row on average
    t["getAcc_g_smth"] = 0.5 * (time.perf_counter() - t0)
    t0 = time.perf_counter()
    O.divcurl_allpairs(pos_p, vel_p, rho_p, m_p, h_p, 0, rows)
    t["div_curlVel"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.accsph_allpairs(pos_p, vel_p, rho_p, P_p, c_p, h_p, m_p, 1.0, dv, cv, 0, rows)
    t["getAcc_sph_shear"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.dudt_allpairs(pos_p, vel_p, rho_p, P_p, c_p, h_p, m_p, 1.0, dv, cv, 0, rows)
    t["get_dU_shear"] = time.perf_counter() - t0
    sample_s = sum(t.values())
    step_s = sample_s * n / rows
    return step_s, sample_s, t


def cpu_tree_oracle(sim_state, m, cfg):
    """The oracle's own FP64 tree path (not the reference's algorithm): tree neighbour
    sets, CSR density, div/curl, acceleration, du/dt and group-walk tree gravity at the
    device state's positions and h (the h solve is not timed: the oracle's Newton is
    all-pairs).  Returns (seconds, per-function seconds)."""
    import oracle as O

    pos = sim_state["pos"]; vel = sim_state["vel"]; h = sim_state["h"]; u = sim_state["u"]
    t = {}
    t0 = time.perf_counter()
    off, idx = O.Tree(pos, m, h=h).neighbours()
    t["tree_neighbours"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    rho = O.density_csr(pos, m, h, off, idx)
    t["density"] = time.perf_counter() - t0
    P = (cfg.gamma - 1) * rho * u
    c = np.sqrt(cfg.gamma * (cfg.gamma - 1) * u)
    t0 = time.perf_counter()
    dv, cv = O.divcurl_csr(pos, vel, rho, m, h, off, idx)
    t["divcurl"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.accsph_csr(pos, vel, rho, P, c, h, m, cfg.alpha, off, idx, dv, cv)
    t["accsph"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.dudt_csr(pos, vel, rho, P, c, h, m, cfg.alpha, off, idx, dv, cv)
    t["dudt"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.Tree(pos, m).gravity(theta=cfg.theta, G=cfg.G, soft_kind=O.SOFT_PLUMMER, eps_plummer=cfg.eps_plummer)
    t["tree_gravity"] = time.perf_counter() - t0
    return sum(t.values()), t


def c1_section(cfg, run):
    """The CPU-runnable Evrard case, N = 10,000: one full step of the FP64 all-pairs
    restatement of the reference step (every row, no sampling), the oracle's tree path,
    and the device step at the same N (launch-bound at this size)."""
    from sph_3_amd import ic

    n = 10000
    d = ic.evrard(n, seed=42)
    sim, dt = run("mixed", 5, 2, d=d)
    g = sim.get()
    del sim
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    step_s, sample_s, parts = cpu_baseline(g, n, d["m"], n)
    tree_s, tree_parts = cpu_tree_oracle(g, d["m"], cfg)
    return {
        "workload": "evrard_collapse, n = 10000", "n": n, "gpu_ms_per_step": 1e3 * dt / 5,
        "gpu_value": n * 5 / dt,
        "cpu_reference_step": {"seconds": step_s, "value": n / step_s, "unit": "particle-steps/s", "cores": threads,
                               "kind": "port", "sample": "all 10000 rows of the all-pairs restatement (a full step)",
                               "function_seconds": parts},
        "cpu_tree_oracle_step": {"seconds": tree_s, "value": n / tree_s, "unit": "particle-steps/s", "cores": threads,
                                 "kind": "oracle tree path (not the reference's algorithm), without the h solve",
                                 "function_seconds": tree_parts},
    }


MLP_FLOP_PER_SAMPLE = 2 * (4 * 128 + 2 * 128 * 128 + 128)  # forward multiply-adds x 2
BF16_MFMA_PEAK_TFLOPS = 2500.0


def traffic_table():
    """HBM bytes per launch measured by rocprofv3 FETCH_SIZE/WRITE_SIZE runs (scaled by
    the copy-kernel calibration), from the JSON block of profiles/hbm_traffic_evrard_1M.md."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_evrard_1M.md")
    try:
        text = open(path).read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        return json.loads(block)
    except (OSError, IndexError, ValueError):
        return {}


def mlp_section(n_inf, epochs=200):
    """Cooling surrogate at the 4M-particle cooling configuration: on-device training
    (2M exact samples), calibration, gated inference throughput and accuracy."""
    from sph_3_amd import cooling as C

    tab = C.CoolingTable()
    net = C.CoolingSurrogate(tab, seed=0)
    mlp_section.net = net  # reused by the cooling step
    n_samples = 2_000_000
    t0 = time.perf_counter()
    losses = net.train(n_samples=n_samples, epochs=epochs)
    t_train = time.perf_counter() - t0
    net.calibrate()
    tol = 0.05
    tm = net.timing(n=n_inf, reps=20, tol=tol)
    fwd = MLP_FLOP_PER_SAMPLE * n_inf / (tm["ms_network"] * 1e-3) / 1e12
    # accuracy on random points of the box against the exact FP64 function
    rng = np.random.default_rng(5)
    lo = np.array([r[0] for r in C.TABLE_RANGES])
    hi = np.array([r[1] for r in C.TABLE_RANGES])
    v = rng.uniform(lo, hi, (200_000, 4))
    _, lam = C.coolingHeatingRates(10 ** v[:, 1], 10 ** v[:, 0], J21=10 ** v[:, 3], Z=10 ** v[:, 2])
    exact = np.log10(np.maximum(lam, 1e-40))
    lg, _, _ = net.rates(*v.T, tol=tol)
    lt, _ = tab.lookup(*v.T)
    rms = lambda e: float(np.sqrt(np.mean(e * e)))  # noqa: E731
    return {
        "workload": f"cooling surrogate 4-128-128-128-1 SiLU, {n_inf} particles uniform in the table box",
        "train": {"samples": n_samples, "epochs": epochs, "seconds": t_train, "final_loss": float(losses[-1]),
                  "tflops": 3 * MLP_FLOP_PER_SAMPLE * n_samples * epochs / t_train / 1e12},
        "ms_gated": tm["ms_gated"], "ms_network": tm["ms_network"], "ms_table_lookup": tm["ms_table"],
        "fallback_fraction": tm["fallback_fraction"], "tol_dex": tol,
        "rms_error_dex": {"gated": rms(lg - exact), "table": rms(lt - exact)},
        "bytes": {"network": net.nbytes, "table_log_lambda": 4 * int(np.prod([r[2] for r in C.TABLE_RANGES]))},
        "roofline": {"kernel": "k_mlp_forward<true>", "bound": "mfma", "achieved": fwd,
                     "peak": BF16_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": fwd / BF16_MFMA_PEAK_TFLOPS,
                     "algorithmic": f"{MLP_FLOP_PER_SAMPLE} flop x {n_inf} samples",
                     "ms_per_launch": tm["ms_network"]},
    }


def cooling_section(args, cfg, run):
    """Evrard collapse plus cooling at N = cooling_n (4,194,304 by default): per-particle
    Z log-uniform in [1e-3, 1] Z_sun and J21 in [1e-2, 1e2], code units M = 1e10 M_sun,
    R = 10 kpc, G = 1; after each step's u update the explicit cooling update with
    log Lambda from the gated surrogate (table fallback at 0.05 dex) and log Gamma and
    T(u) from tables.  The all-FP64 form takes the table for log Lambda (there is no
    FP64 network).  CPU: the same all-pairs restatement on a quarter of the 1M row
    sample (rows cost 4x at 4x N) plus the oracle's exact cooling update."""
    from sph_3_amd import cooling as C
    from sph_3_amd import ic

    n = args.cooling_n
    d = ic.evrard(n, seed=44)
    Z, J = ic.cooling_fields(n, seed=45)
    net = getattr(mlp_section, "net", None)
    mode = "mlp" if net is not None else "table"
    units = C.CoolingParams.from_mass_length(1e10 * C.M_SUN_G, 10 * C.KPC_CM, mode=mode, tol=0.05)

    def cooling(mode_):
        def f(sim):
            u = C.CoolingParams(**{**units.__dict__, "mode": mode_})
            sim.set_cooling(Z, J, u, surrogate=net if mode_ == "mlp" else None)
        return f

    sim, dt = run("mixed", args.steps, args.warmup, d=d, cooling=cooling(mode))
    st = sim.stats()
    g = sim.get()
    ms = 1e3 * dt / args.steps
    per_step = {k: st["phase_ms"][k] / args.steps for k in st["phase_ms"] if st["phase_calls"][k]}
    res = {
        "workload": f"evrard_collapse + cooling, n = {n}", "n": n, "cooling_mode": mode, "tol_dex": units.tol,
        "units": {"mass_g": 1e10 * C.M_SUN_G, "length_cm": 10 * C.KPC_CM, "density_cgs": units.unit_density_cgs,
                  "u_cgs": units.unit_u_cgs, "time_s": units.unit_time_s},
        "ms_per_step": ms, "value": n * args.steps / dt, "unit": "particle-steps/s",
        "phase_ms_per_step": per_step,
        "cooling_fallback_fraction_last_step": st["cool_fallback"] / n,
        "cooling_floored_fraction_last_step": st["cool_floor"] / n,
    }
    del sim
    if not args.no_fp64_form:
        steps64 = max(2, args.steps // 2)
        sim64, dt64 = run("fp64", steps64, 1, d=d, cooling=cooling("table"))
        res["fp64_form"] = {"ms_per_step": 1e3 * dt64 / steps64, "mixed_speedup": (1e3 * dt64 / steps64) / ms,
                            "cooling_mode": "table"}
        del sim64
    if not args.no_cpu_baseline:
        import oracle as O

        rows = max(64, args.cpu_rows // 4)
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log(f"[bench] cooling workload cpu baseline: {rows} rows x {n} sources on {threads} threads")
        step_s, sample_s, parts = cpu_baseline(g, n, d["m"], rows)
        sel = np.random.default_rng(9).choice(n, rows, replace=False)
        t0 = time.perf_counter()
        box = C.TABLE_RANGES
        O.cool_update(g["rho"][sel], g["u"][sel], Z[sel], J[sel], cfg.dt,
                      (units.unit_density_cgs, units.unit_u_cgs, units.unit_time_s, units.XH, units.u_floor_cgs),
                      [b[0] for b in box], [b[1] for b in box])
        t_cool = time.perf_counter() - t0
        parts["cooling_update"] = t_cool
        step_s += t_cool * n / rows
        res["cpu_baseline"] = {
            "value": n / step_s, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"{rows} random target rows x all {n} sources of the FP64 all-pairs restatement of the "
                      f"step plus the exact FP64 cooling update of {rows} particles; "
                      f"{sample_s + t_cool:.1f} s measured, extrapolated x{n / rows:.0f}",
            "step_seconds_extrapolated": step_s, "function_seconds_sample": parts,
        }
        res["speedup_vs_cpu"] = res["value"] / res["cpu_baseline"]["value"]
    return res


def agn_section(args, cfg, run):
    """AGN-outflow-like workload: a lognormal clumpy cloud (sigma_ln rho = 1.5, Gaussian
    field on 256^3 cells) of agn_cloud particles at T = 1e4 K plus agn_wind hot (1e7 K)
    wind particles leaving r = 0.05 R at v = 0.1; gravity, SPH and the gated-surrogate
    cooling (table fallback at 0.05 dex); dt = 1e-5 (the wind shell's Courant limit); the
    outflow driver's source terms: an isothermal halo (sigma = 2, about 130 km/s in these
    units) and 50.5 outflow pairs per step on average injected within 1 pc (1e-4) of the
    origin at v = 30 (about 2,000 km/s), T = 1e4 K."""
    from sph_3_amd import cooling as C
    from sph_3_amd import ic
    from sph_3_amd.step import AgnParams

    net = getattr(mlp_section, "net", None)
    mode = "mlp" if net is not None else "table"
    units = C.CoolingParams.from_mass_length(1e10 * C.M_SUN_G, 10 * C.KPC_CM, mode=mode, tol=0.05)
    t0 = time.perf_counter()
    d = ic.agn_cloud(args.agn_cloud, args.agn_wind, n_grid=256, u_cloud=ic.u_of_T(1e4, units.unit_u_cgs),
                     u_wind=ic.u_of_T(1e7, units.unit_u_cgs))
    n = len(d["m"])
    Z, J = ic.cooling_fields(n, seed=49)
    log(f"[bench] agn workload: {n} particles generated in {time.perf_counter() - t0:.1f} s")
    dt_saved = cfg.dt
    cfg.dt = 1e-5

    m_inj = float(d["m"][0])
    agn = AgnParams(sigma=2.0, M_dot=50.5 * 2 * m_inj / 1e-5, v_in=30.0, m_inject=m_inj,
                    u_inject=ic.u_of_T(1e4, units.unit_u_cgs), r_inject=1e-4, Z_inject=1.0, J21_inject=1.0,
                    seed=2025, h_ref_index=n - 1)
    n_max = n + 2 * 51 * (args.steps + args.warmup + 4)

    def cooling(mode_):
        def f(sim):
            u = C.CoolingParams(**{**units.__dict__, "mode": mode_})
            sim.set_cooling(Z, J, u, surrogate=net if mode_ == "mlp" else None)
            sim.set_agn(agn)
        return f

    try:
        sim, dt = run("mixed", args.steps, args.warmup, d=d, cooling=cooling(mode), n_max=n_max)
        st = sim.stats()
        ms = 1e3 * dt / args.steps
        n_end = sim.n
        res = {
            "workload": f"agn-like: {args.agn_cloud} cloud + {args.agn_wind} wind particles", "n": n,
            "cooling_mode": mode, "dt": cfg.dt, "ms_per_step": ms, "value": n * args.steps / dt,
            "unit": "particle-steps/s",
            "phase_ms_per_step": {k: st["phase_ms"][k] / args.steps for k in st["phase_ms"] if st["phase_calls"][k]},
            "neighbours_per_particle": st["ngb_total"] / n, "neighbour_layout": "padded rows or CSR fallback",
            "gravity_interactions_per_particle": (st["grav_pp"] + st["grav_pn"]) / n,
            "h_solve_rounds_last_step": st["h_rounds"],
            "h_solve_walk_solved_targets_last_step": st["h_big"],
            "cooling_fallback_fraction_last_step": st["cool_fallback"] / n,
            "source_terms": {"halo_sigma": agn.sigma, "pairs_per_step": "50 or 51 (50.5 on average)",
                             "particles_injected": n_end - n},
        }
        del sim
        if not args.no_fp64_form:
            steps64 = max(2, args.steps // 2)
            sim64, dt64 = run("fp64", steps64, 1, d=d, cooling=cooling("table"), n_max=n_max)
            res["fp64_form"] = {"ms_per_step": 1e3 * dt64 / steps64, "mixed_speedup": (1e3 * dt64 / steps64) / ms,
                                "cooling_mode": "table"}
            del sim64
    finally:
        cfg.dt = dt_saved
    return res


def sedov_section(args, cfg, run):
    """Periodic 3D Sedov blast: sedov_side^3 particles on a jittered lattice (+-0.1
    spacing) in the unit periodic box, rho = 1, u = 1e-5, E = 1 deposited
    kernel-weighted on the 64 particles nearest the centre, no gravity; minimum-image
    neighbour search and hydro, dt = 5e-6 (the blast's Courant limit)."""
    from sph_3_amd import ic

    t0 = time.perf_counter()
    d = ic.sedov(args.sedov_side)
    n = len(d["m"])
    log(f"[bench] sedov: {n} particles generated in {time.perf_counter() - t0:.1f} s")
    saved = (cfg.dt, cfg.G, cfg.period)
    cfg.dt, cfg.G, cfg.period = 5e-6, 0.0, 1.0
    try:
        sim, dt = run("mixed", args.steps, args.warmup, d=d)
        st = sim.stats()
        ms = 1e3 * dt / args.steps
        res = {
            "workload": f"sedov blast, periodic unit box, {args.sedov_side}^3 = {n} particles", "n": n,
            "dt": cfg.dt, "ms_per_step": ms, "value": n * args.steps / dt, "unit": "particle-steps/s",
            "phase_ms_per_step": {k: st["phase_ms"][k] / args.steps for k in st["phase_ms"] if st["phase_calls"][k]},
            "neighbours_per_particle": st["ngb_total"] / n, "h_solve_rounds_last_step": st["h_rounds"],
            "h_solve_walk_solved_targets_last_step": st["h_big"],
        }
        del sim
        if not args.no_fp64_form:
            steps64 = max(2, args.steps // 2)
            sim64, dt64 = run("fp64", steps64, 1, d=d)
            res["fp64_form"] = {"ms_per_step": 1e3 * dt64 / steps64, "mixed_speedup": (1e3 * dt64 / steps64) / ms}
            del sim64
    finally:
        cfg.dt, cfg.G, cfg.period = saved
    return res


def column_section(args, cfg, run, ms_step, kbar):
    """Column density N_H from |grad rho| (the CLOUDY driver's grad_rho_norm_ngb and N_H
    loop) and the per-particle time-step criteria (dt_array_indiv_dt, C_CFL = 0.3) computed
    in every force evaluation of the 1M Evrard step, units 1e10 M_sun and 10 kpc: each pass's
    HIP-event time, the column pass's algorithmic bandwidth and the step time with both
    against the headline step."""
    from sph_3_amd import column as C

    units = C.ColumnParams.from_mass_length(1.989e43, 3.0857e22)
    cfg.C_CFL = 0.3
    try:
        sim, dt = run("mixed", args.steps, args.warmup, cooling=lambda s: s.set_column(units))
    finally:
        cfg.C_CFL = 0.0
    st = sim.stats()
    n = st["n"]
    ms_col = st["phase_ms"]["column"] / max(1, st["phase_calls"]["column"])
    ms_ts = st["phase_ms"]["timestep"] / max(1, st["phase_calls"]["timestep"])
    lognh, _ = sim.column()
    dt_min = st["dt_min"]
    del sim
    bpp = 4.0 * kbar + 8.0 + 24 + 8 + 8 + 4 + 16  # row, x, h, m, rho in; |grad rho|, log N_H out
    return {
        "workload": f"evrard_collapse + column density, n = {n}", "n": n,
        "ms_column_pass": ms_col,
        "pass_bandwidth": {"GB/s": bpp * n / (ms_col * 1e-3) / 1e9,
                           "frac": bpp * n / (ms_col * 1e-3) / 1e9 / HBM_PEAK_GBS, "bytes_per_particle": bpp},
        "ms_timestep_pass": ms_ts, "dt_min_C_CFL_0.3": dt_min,
        "ms_per_step_with_both": 1e3 * dt / args.steps, "ms_per_step_without": ms_step,
        "log_NH_range_cm2": [float(np.min(lognh)), float(np.median(lognh)), float(np.max(lognh))],
    }


def io_section(g, m):
    """Host file I/O of the 1M state in the reference's snapshot layout (saveArraysToBinary)
    and IC layout (readVectorsFromFile), through the page cache of the box's /tmp."""
    import tempfile

    from sph_3_amd import snapshot as S

    d = S.state_arrays(g["pos"], g["vel"], m, g["u"], g["h"], g["rho"])
    n = len(m)
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, write, read, nbytes in [
                ("snapshot", lambda f: S.write_snapshot(f, d), S.read_snapshot, 4 + 44 * n),
                ("ic_header", lambda f: S.write_ic(f, d, layout="header"), lambda f: S.read_ic(f, "header"),
                 4 + 44 * n)]:
            fn = os.path.join(tmp, name + ".bin")
            t0 = time.perf_counter(); write(fn); tw = time.perf_counter() - t0
            t0 = time.perf_counter(); read(fn); tr = time.perf_counter() - t0
            out[name] = {"bytes": nbytes, "write_GB/s": nbytes / tw / 1e9, "read_GB/s": nbytes / tr / 1e9}
    return out


def copy_probe(nbytes):
    """A plain device copy of known size (the calibration run of the HBM traffic counters)."""
    import torch

    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes, "launches": 3}), flush=True)


def accel_error(d, cfg):
    """Acceleration of the mixed form against the FP64 form after the first force evaluation
    from the same initial state (the h solve is FP64 in both, so the tree, h and neighbour
    sets agree and the difference is the FP32 pair and gravity arithmetic)."""
    from sph_3_amd.step import DeviceSPH

    acc = {}
    for prec in ("fp64", "mixed"):
        cfg.precision = prec
        sim = DeviceSPH(d["pos"], d["vel"], d["m"], d["u"], d["h0"], cfg)
        sim.prime()
        acc[prec] = sim.get()["acc"]
        del sim
    a64, am = acc["fp64"], acc["mixed"]
    na = np.linalg.norm(a64, axis=1)
    de = np.linalg.norm(am - a64, axis=1)
    rel = de / np.maximum(na, np.finfo(float).tiny)
    return {"definition": "|a_mixed - a_fp64| / |a_fp64| per particle, first force evaluation",
            "max": float(rel.max()), "p99.9": float(np.quantile(rel, 0.999)), "rms": float(np.sqrt(np.mean(rel ** 2))),
            "max_over_rms_accel": float(de.max() / np.sqrt(np.mean(na ** 2))), "bound": 1e-3,
            "within_bound": bool(rel.max() <= 1e-3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1048576)
    ap.add_argument("--cpu-rows", type=int, default=4096)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64-form", action="store_true")
    ap.add_argument("--no-mlp", action="store_true")
    ap.add_argument("--mlp-n", type=int, default=4194304)
    ap.add_argument("--mlp-epochs", type=int, default=200, help="training passes of the surrogate")
    ap.add_argument("--no-cooling-step", action="store_true")
    ap.add_argument("--cooling-n", type=int, default=4194304)
    ap.add_argument("--no-agn", action="store_true")
    ap.add_argument("--no-sedov", action="store_true")
    ap.add_argument("--no-column", action="store_true", help="skip the column-density measurement")
    ap.add_argument("--sedov-side", type=int, default=128)
    ap.add_argument("--agn-cloud", type=int, default=7340032)
    ap.add_argument("--agn-wind", type=int, default=1048576)
    ap.add_argument("--copy-probe", type=int, default=0, help="only run a device copy of this many bytes")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); sharding the particle "
              "set over several GPUs is designed but not built or measured", file=sys.stderr)
        sys.exit(2)

    from sph_3_amd import ic
    from sph_3_amd.step import DeviceSPH, StepConfig

    n = args.n
    d = ic.evrard(n, seed=42)
    cfg = StepConfig(dt=1e-3, gamma=5 / 3, alpha=1.0, G=1.0, theta=0.5, softening="plummer", eps_plummer=1e-3,
                     n_ngb=64.0, h_tol=1e-10, precision="mixed", balsara=True)

    def run(precision, steps, warmup, d=d, cooling=None, n_max=None):
        cfg.precision = precision
        sim = DeviceSPH(d["pos"], d["vel"], d["m"], d["u"], d["h0"], cfg, n_max=n_max)
        if cooling is not None:
            cooling(sim)
        log(f"[bench] {precision}: prime")
        sim.prime()
        for w in range(warmup):
            sim.step()
            log(f"[bench] {precision}: warmup step {w + 1}/{warmup} done")
        sim.reset_stats()
        try:
            import torch

            torch.cuda.synchronize()
        except Exception:
            pass
        t0 = time.perf_counter()
        for k in range(steps):
            sim.step(timing=True)
        try:
            import torch

            torch.cuda.synchronize()
        except Exception:
            pass
        dt = time.perf_counter() - t0
        log(f"[bench] {precision}: {steps} timed steps in {dt:.3f} s")
        return sim, dt

    sim, dt = run("mixed", args.steps, args.warmup)
    st = sim.stats()
    g_mixed = sim.get()  # the context is shared: read the state before the next run replaces it
    ms_step = 1e3 * dt / args.steps
    value = n * args.steps / dt
    phases = {k: v / max(1, st["phase_calls"][k]) for k, v in st["phase_ms"].items() if st["phase_calls"][k]}
    per_step = {k: st["phase_ms"][k] / args.steps for k in phases}
    kbar = st["ngb_total"] / n
    bytes_pp = per_particle_bytes(kbar)
    bytes_pp["moments_h"] = bytes_pp["moments"]
    # dominant kernel: the phase with the most time per step
    dom = max(per_step, key=per_step.get)
    h_flops = H_FLOP_PER_TEST * st["h_tests"] + H_FLOP_PER_TERM * st["h_terms"]
    if dom == "hsolve":
        ach = h_flops / (phases[dom] * 1e-3) / 1e12
        roof = {"kernel": "k_hsolve", "bound": "valu_fp64", "achieved": ach, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": ach / FP64_PEAK_TFLOPS, "traffic": None,
                "algorithmic": f"{H_FLOP_PER_TEST} flop x {st['h_tests']} walk distance tests + "
                               f"{H_FLOP_PER_TERM} flop x {st['h_terms']} Newton kernel terms",
                "ms_per_launch": phases[dom]}
    elif dom == "gravity":
        flops = FLOP_PER_INTERACTION * (st["grav_pp"] + st["grav_pn"])
        ach = flops / (phases[dom] * 1e-3) / 1e12
        roof = {"kernel": KERNEL_OF_PHASE[dom], "bound": "valu", "achieved": ach, "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": ach / FP32_PEAK_TFLOPS, "traffic": None,
                "algorithmic": f"{FLOP_PER_INTERACTION} flop x {st['grav_pp'] + st['grav_pn']} interactions",
                "ms_per_launch": phases[dom],
                # the 157.3 TFLOP/s peak counts packed FP32 (two FMAs per lane and cycle), which the
                # build disables; the non-packed VALU ceiling is half of it
                "peak_nonpacked": FP32_PEAK_TFLOPS / 2, "frac_of_nonpacked": ach / (FP32_PEAK_TFLOPS / 2)}
    else:
        b = bytes_pp.get(dom, 0.0)
        ach = b * n / (phases[dom] * 1e-3) / 1e9
        roof = {"kernel": KERNEL_OF_PHASE.get(dom, dom), "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": None,
                "algorithmic": f"{b:.1f} B/particle x {n} particles", "ms_per_launch": phases[dom]}
    # HBM fractions of the CSR passes (reported beside the dominant kernel)
    hbm = {k: {"ms": phases[k], "GB/s": bytes_pp[k] * n / (phases[k] * 1e-3) / 1e9,
               "frac": bytes_pp[k] * n / (phases[k] * 1e-3) / 1e9 / HBM_PEAK_GBS}
           for k in bytes_pp if k in phases and not (k == "moments" and cfg.softening == "plummer")}
    grav_tflops = None
    if "gravity" in phases:
        grav_tflops = FLOP_PER_INTERACTION * (st["grav_pp"] + st["grav_pn"]) / (phases["gravity"] * 1e-3) / 1e12
    out = {
        "metric": "SPH particle-steps/sec on 1 GPU, mixed-precision vs FP64, at <=1e-3 accel error",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32",
        "precision": "mixed: f32 pair/gravity arithmetic with Kahan f32 sums; f64 positions, velocities, "
                     "centres of mass, h solve, reductions",
        "data": "synthetic (seeded Evrard sphere)",
        "config": {"workload": "evrard_collapse", "n": n, "M": 1.0, "R": 1.0, "G": 1.0, "u0": 0.05,
                   "gamma": 5 / 3, "n_ngb": 64, "theta": 0.5, "softening": "plummer", "eps": 1e-3,
                   "alpha": 1.0, "balsara": True, "dt": 1e-3},
        "phases_ms": phases,
        "phase_ms_per_step": per_step,
        "h_solve_fp64_tflops": h_flops / (phases["hsolve"] * 1e-3) / 1e12 if "hsolve" in phases else None,
        "neighbours_per_particle": kbar,
        "neighbour_pairs_from_other_side": st["ngb_extra"],
        "h_solve_rounds_last_step": st["h_rounds"],
        "h_solve_walk_solved_targets_last_step": st["h_big"],
        "gravity_interactions_per_particle": (st["grav_pp"] + st["grav_pn"]) / n,
        "gravity_tflops": grav_tflops,
        "pass_bandwidth": hbm,
        "roofline": roof,
    }
    tr = traffic_table().get(roof["kernel"]) if n == 1048576 else None  # measured at this N only
    if tr:
        roof["traffic"] = tr["fetch_bytes"] + tr["write_bytes"]
        roof["traffic_source"] = "profiles/hbm_traffic_evrard_1M.md"
    if not args.no_mlp:
        log("[bench] cooling surrogate: train, calibrate, time")
        out["mlp_cooling"] = mlp_section(args.mlp_n, args.mlp_epochs)
    if not args.no_fp64_form:
        sim64, dt64 = run("fp64", max(2, args.steps // 2), 1)
        out["fp64_form"] = {"ms_per_step": 1e3 * dt64 / max(2, args.steps // 2),
                            "mixed_speedup": (1e3 * dt64 / max(2, args.steps // 2)) / ms_step}
        del sim64
        log("[bench] acceleration error of the mixed form against the FP64 form")
        out["fp64_form"]["accel_error"] = accel_error(d, cfg)
    if not args.no_cpu_baseline:
        g = g_mixed
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log(f"[bench] cpu baseline: {args.cpu_rows} rows x {n} sources on {threads} threads")
        step_s, sample_s, parts = cpu_baseline(g, n, d["m"], args.cpu_rows)
        out["cpu_baseline"] = {
            "value": n / step_s, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"{args.cpu_rows} random target rows x all {n} sources of the FP64 all-pairs restatement "
                      f"of sph_serial_v2.py's step (h_smooth_fast, getDensity, getAcc_g_smth, div_curlVel, "
                      f"getAcc_sph_shear, get_dU_shear); {sample_s:.1f} s measured, extrapolated x{n / args.cpu_rows:.0f}",
            "step_seconds_extrapolated": step_s,
            "function_seconds_sample": parts,
        }
        out["speedup_vs_cpu"] = value / out["cpu_baseline"]["value"]
        log(f"[bench] cpu tree oracle at n = {n} on {threads} threads")
        tree_s, tree_parts = cpu_tree_oracle(g, d["m"], cfg)
        out["cpu_tree_oracle"] = {
            "value": n / tree_s, "unit": "particle-steps/s", "cores": threads,
            "kind": "oracle tree path (not the reference's algorithm): tree neighbours, CSR hydro, group-walk "
                    "gravity at the device's h, without the h solve; every row",
            "seconds": tree_s, "function_seconds": tree_parts,
        }
        log("[bench] the cpu-runnable case, n = 10000")
        out["evrard_10k"] = c1_section(cfg, run)
    out["file_io"] = io_section(g_mixed, d["m"])
    if not args.no_column:
        log("[bench] column density and time-step criteria in the step")
        out["column_density"] = column_section(args, cfg, run, ms_step, kbar)
    if not args.no_cooling_step:
        out["evrard_cooling"] = cooling_section(args, cfg, run)
    if not args.no_agn:
        out["agn_outflow"] = agn_section(args, cfg, run)
    if not args.no_sedov:
        out["sedov_blast"] = sedov_section(args, cfg, run)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
