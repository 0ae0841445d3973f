"""Throughput benchmark of the SPH step on one MI355X.

Headline workload (``value``): Evrard collapse plus cooling, N = 4,194,304
equal-mass gas particles (M = 1, R = 1, G = 1 in code units of 1e10 M_sun and
10 kpc, rho ~ 1/r from r = R sqrt(U), seeded isotropic directions, u = 0.05,
gamma = 5/3), cubic spline, N_ngb = 64 (FP64 Newton h solve), theta = 0.5,
Plummer eps = 1e-3, alpha = 1 with the Balsara switch, dt = 1e-3, per-particle
Z and J21, and after each u update the explicit cooling update with log Lambda
from the BF16 surrogate (table fallback).  One "step" is the full KDK step of
sph_serial_v2.py on the device: kick/drift, Morton sort, tree, h solve,
neighbour rows, density+EOS+div/curl, tree gravity, acceleration, kick,
du/dt, u update, cooling.

Prints ONE JSON line: particle-steps/s (whole job, inputs resident in HBM),
the roofline of the dominant kernel from HIP events on the library's stream
(HBM traffic per launch from the committed PMC summary at that N), the
all-FP64 form of the same kernels with the acceleration error between the
forms, the CPU baseline (the oracle's FP64 all-pairs restatement of the
reference step on 4,096 sampled rows, with its sampling spread and the host's
nproc/lscpu), and the other configurations as keys: the surrogate's training
and inference (``mlp_cooling``), the 1,048,576-particle Evrard step with its
errors against the FP64 oracle (``evrard_1m``), the 10,000-particle
CPU-runnable case (``evrard_10k``), the AGN-like 8.4M workload and the
periodic Sedov blast, each with its own roofline.
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
                   "gravity": "k_gravity_group<float,2,false>", "moments_h": "k_moments_local+k_moments_up"}


def host_info():
    """nproc and the lscpu lines that describe the host's cores."""
    import subprocess

    info = {"nproc": None, "affinity_cpus": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}
    try:
        info["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True, timeout=10).stdout.strip())
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        keep = ("Model name", "CPU(s)", "Thread(s) per core", "Core(s) per socket", "Socket(s)", "CPU max MHz",
                "NUMA node(s)", "L3 cache")
        info["lscpu"] = {k.strip(): v.strip() for k, v in (ln.split(":", 1) for ln in out.splitlines() if ":" in ln)
                         if k.strip() in keep}
    except (OSError, subprocess.SubprocessError):
        info["lscpu"] = None
    return info


def cpu_baseline(sim_state, n, m, rows, seed=0, batches=8):
    """Time the oracle's FP64 all-pairs restatement of the reference step
    (libsx.h_smooth_fast, getDensity, getAcc_g_smth, shear.div_curlVel,
    getAcc_sph_shear, get_dU_shear) on `rows` sampled target rows against all
    n sources; extrapolate linearly in rows (rows are independent).  Each function
    runs over `batches` consecutive batches of the sample, whose spread gives the
    relative standard error of its extrapolated time."""
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
    h_start = h_p * (50.0 / 64.0) ** (1 / 3)
    funcs = {
        "h_smooth_fast": lambda a, b: O.h_smooth_fast(pos_p, h_start, nbeg=a, nend=b),
        "getDensity": lambda a, b: O.density_allpairs(pos_p, m_p, h_p, a, b),
        # the reference's symmetric j > i loop does half the pairs per row on average: timed x 1/2
        "getAcc_g_smth": lambda a, b: O.grav_direct(pos_p, m_p, eps, 1.0, O.SOFT_SPLINE, 0.0, a, b),
        "div_curlVel": lambda a, b: O.divcurl_allpairs(pos_p, vel_p, rho_p, m_p, h_p, a, b),
        "getAcc_sph_shear": lambda a, b: O.accsph_allpairs(pos_p, vel_p, rho_p, P_p, c_p, h_p, m_p, 1.0, dv, cv,
                                                           a, b),
        "get_dU_shear": lambda a, b: O.dudt_allpairs(pos_p, vel_p, rho_p, P_p, c_p, h_p, m_p, 1.0, dv, cv, a, b),
    }
    edges = np.linspace(0, rows, batches + 1).astype(int)
    t, rse = {}, {}
    var = 0.0
    for name, f in funcs.items():
        bt = []
        for a, b in zip(edges[:-1], edges[1:]):
            t0 = time.perf_counter()
            f(int(a), int(b))
            bt.append((time.perf_counter() - t0) / (b - a))  # seconds per row
        bt = np.array(bt) * (0.5 if name == "getAcc_g_smth" else 1.0)
        t[name] = float(bt.mean() * rows)
        se = float(bt.std(ddof=1) / np.sqrt(len(bt))) if len(bt) > 1 else 0.0
        rse[name] = se / float(bt.mean())
        var += (se * n) ** 2
    sample_s = sum(t.values())
    step_s = sample_s * n / rows
    return step_s, sample_s, t, {"per_function_rel_std_error": rse, "step_rel_std_error": float(np.sqrt(var)) / step_s,
                                 "batches": batches}


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
    a_h = O.accsph_csr(pos, vel, rho, P, c, h, m, cfg.alpha, off, idx, dv, cv)
    t["accsph"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    du = O.dudt_csr(pos, vel, rho, P, c, h, m, cfg.alpha, off, idx, dv, cv)
    t["dudt"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    a_g, cnt = O.Tree(pos, m).gravity(theta=cfg.theta, G=cfg.G, soft_kind=O.SOFT_PLUMMER, eps_plummer=cfg.eps_plummer)
    t["tree_gravity"] = time.perf_counter() - t0
    return sum(t.values()), t, dict(off=off, idx=idx, rho=rho, acc=a_h + a_g, du=du, nint=cnt)


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
    step_s, sample_s, parts, _ = cpu_baseline(g, n, d["m"], n, batches=1)
    tree_s, tree_parts, _ = cpu_tree_oracle(g, d["m"], cfg)
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


TRAFFIC_PROFILES = {1048576: "profiles/hbm_traffic_evrard_1M.md",
                    4194304: "profiles/hbm_traffic_evrard_cooling_4M.md"}


def traffic_table(n):
    """HBM bytes per launch measured by rocprofv3 FETCH_SIZE/WRITE_SIZE runs (scaled by
    the copy-kernel calibration) at this N, from the JSON block of the committed profile
    summary ({} when there is none for this N)."""
    rel = TRAFFIC_PROFILES.get(n)
    if rel is None:
        return {}, None
    try:
        text = open(os.path.join(ROOT, rel)).read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        return json.loads(block), rel
    except (OSError, IndexError, ValueError):
        return {}, None


def roofline(st, n, softening="plummer"):
    """Roofline of the step's dominant phase (most HIP-event time per step, timed steps only)
    and the algorithmic bandwidth of the HBM-bound passes.  Gravity: FP32 VALU, 20 flop per
    counted interaction; h solve: FP64 VALU from its device counters; the rest: HBM bytes."""
    phases = {k: v / max(1, st["phase_calls"][k]) for k, v in st["phase_ms"].items() if st["phase_calls"][k]}
    per_step = {k: st["phase_ms"][k] / max(1, st["steps"]) for k in phases}
    kbar = st["ngb_total"] / n
    bytes_pp = per_particle_bytes(kbar)
    bytes_pp["moments_h"] = bytes_pp["moments"]
    dom = max(per_step, key=per_step.get)
    h_flops = H_FLOP_PER_TEST * st["h_tests"] + H_FLOP_PER_TERM * st["h_terms"]
    if dom == "hsolve":
        ach = h_flops / (phases[dom] * 1e-3) / 1e12
        roof = {"kernel": "k_hwalk+k_hnewton", "bound": "valu_fp64", "achieved": ach, "peak": FP64_PEAK_TFLOPS,
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
    table, src = traffic_table(n)
    tr = table.get(roof["kernel"])
    if tr:
        roof["traffic"] = tr["fetch_bytes"] + tr["write_bytes"]
        roof["traffic_source"] = src
    hbm = {k: {"ms": phases[k], "GB/s": bytes_pp[k] * n / (phases[k] * 1e-3) / 1e9,
               "frac": bytes_pp[k] * n / (phases[k] * 1e-3) / 1e9 / HBM_PEAK_GBS}
           for k in bytes_pp if k in phases and not (k == "moments" and softening == "plummer")}
    extra = {
        "phases_ms": phases, "phase_ms_per_step": per_step,
        "h_solve_fp64_tflops": h_flops / (phases["hsolve"] * 1e-3) / 1e12 if "hsolve" in phases else None,
        "neighbours_per_particle": kbar, "neighbour_pairs_from_other_side": st["ngb_extra"],
        "h_solve_rounds_last_step": st["h_rounds"], "h_solve_walk_solved_targets_last_step": st["h_big"],
        "gravity_interactions_per_particle": (st["grav_pp"] + st["grav_pn"]) / n,
        "gravity_tflops": (FLOP_PER_INTERACTION * (st["grav_pp"] + st["grav_pn"]) / (phases["gravity"] * 1e-3) / 1e12
                           if "gravity" in phases else None),
        "pass_bandwidth": hbm,
    }
    return roof, extra


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
        "workload": f"cooling surrogate 4-128-128-128-1 ReLU, {n_inf} particles uniform in the table box",
        "train": {"samples": n_samples, "epochs": epochs, "seconds": t_train, "final_loss": float(losses[-1]),
                  "tflops": 3 * MLP_FLOP_PER_SAMPLE * n_samples * epochs / t_train / 1e12},
        "ms_gated": tm["ms_gated"], "ms_network": tm["ms_network"], "ms_table_lookup": tm["ms_table"],
        "fallback_fraction": tm["fallback_fraction"], "tol_dex": tol,
        "rms_error_dex": {"gated": rms(lg - exact), "table": rms(lt - exact)},
        "bytes": {"network": net.nbytes, "table_log_lambda": 4 * int(np.prod([r[2] for r in C.TABLE_RANGES]))},
        "roofline": {"kernel": "k_mlp_forward", "bound": "mfma", "achieved": fwd,
                     "peak": BF16_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": fwd / BF16_MFMA_PEAK_TFLOPS,
                     "algorithmic": f"{MLP_FLOP_PER_SAMPLE} flop x {n_inf} samples",
                     "ms_per_launch": tm["ms_network"]},
    }


def cooling_setup(args):
    """The Evrard-plus-cooling workload at N = cooling_n (4,194,304 by default): seeded
    Evrard sphere, per-particle Z log-uniform in [1e-3, 1] Z_sun and J21 in [1e-2, 1e2],
    code units M = 1e10 M_sun, R = 10 kpc, G = 1.  Returns (d, Z, J21, units, cooling)
    with cooling(mode) a set-up callback: after each step's u update the explicit cooling
    update with log Lambda from the gated surrogate ("mlp", table fallback at 0.05 dex)
    or the table ("table"), log Gamma and T(u) from tables."""
    from sph_3_amd import cooling as C
    from sph_3_amd import ic

    n = args.cooling_n
    d = ic.evrard(n, seed=44)
    Z, J = ic.cooling_fields(n, seed=45)
    net = getattr(mlp_section, "net", None)
    units = C.CoolingParams.from_mass_length(1e10 * C.M_SUN_G, 10 * C.KPC_CM, mode="mlp" if net else "table", tol=0.05)

    def cooling(mode_):
        def f(sim):
            u = C.CoolingParams(**{**units.__dict__, "mode": mode_})
            sim.set_cooling(Z, J, u, surrogate=net if mode_ == "mlp" else None)
        return f

    return d, Z, J, units, cooling


def prime_states(d, cfg, cooling=None, neighbours=False):
    """The first force evaluation of both forms from the same initial state (the h solve is
    FP64 in both, so the tree, h and neighbour sets agree and the forms differ by the FP32
    pair and gravity arithmetic).  Returns {form: state} (+ the mixed form's sets)."""
    from sph_3_amd.step import DeviceSPH

    out = {}
    saved = cfg.precision
    try:
        for prec in ("fp64", "mixed"):
            cfg.precision = prec
            sim = DeviceSPH(d["pos"], d["vel"], d["m"], d["u"], d["h0"], cfg)
            if cooling is not None:
                cooling(sim)
            sim.prime()
            out[prec] = sim.get()
            if neighbours and prec == "mixed":
                out["sets"] = sim.neighbours()
            del sim
    finally:
        cfg.precision = saved
    return out


def accel_error(states):
    a64, am = states["fp64"]["acc"], states["mixed"]["acc"]
    na = np.linalg.norm(a64, axis=1)
    de = np.linalg.norm(am - a64, axis=1)
    rel = de / np.maximum(na, np.finfo(float).tiny)
    return {"definition": "|a_mixed - a_fp64| / |a_fp64| per particle, first force evaluation",
            "max": float(rel.max()), "p99.9": float(np.quantile(rel, 0.999)), "rms": float(np.sqrt(np.mean(rel ** 2))),
            "max_over_rms_accel": float(de.max() / np.sqrt(np.mean(na ** 2))), "bound": 1e-3,
            "within_bound": bool(rel.max() <= 1e-3)}


def oracle_errors(states, ref, d, dt):
    """Errors of the device's first force evaluation against the FP64 oracle evaluated at the
    device's positions and h: rho relative, acceleration and du/dt per particle relative
    (|x - y| / |y|) and normalised (|x - y| / rms |y|), max and rms."""
    out = {}
    for prec in ("mixed", "fp64"):
        g = states[prec]
        e_rho = np.abs(g["rho"] / ref["rho"] - 1)
        da = np.linalg.norm(g["acc"] - ref["acc"], axis=1)
        na = np.linalg.norm(ref["acc"], axis=1)
        du = (g["u"] - d["u"]) / dt
        ddu = np.abs(du - ref["du"])
        rms_a, rms_du = np.sqrt(np.mean(na ** 2)), np.sqrt(np.mean(ref["du"] ** 2))
        out[prec] = {
            "rho_rel": {"max": float(e_rho.max()), "rms": float(np.sqrt(np.mean(e_rho ** 2)))},
            "acc_rel": {"max": float((da / na).max()), "p99.9": float(np.quantile(da / na, 0.999)),
                        "rms": float(np.sqrt(np.mean((da / na) ** 2)))},
            "acc_normalised": {"max": float(da.max() / rms_a), "rms": float(np.sqrt(np.mean(da ** 2)) / rms_a)},
            "dudt_normalised": {"max": float(ddu.max() / rms_du), "rms": float(np.sqrt(np.mean(ddu ** 2)) / rms_du)},
        }
    return out


def cooling_cpu_baseline(args, cfg, g, d, Z, J, units):
    """CPU baseline of the cooling workload: the all-pairs restatement on cpu_rows sampled
    rows x all N sources plus the oracle's exact FP64 cooling update of as many particles."""
    import oracle as O
    from sph_3_amd import cooling as C

    n = len(d["m"])
    rows = args.cpu_rows
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    log(f"[bench] cooling workload cpu baseline: {rows} rows x {n} sources on {threads} threads")
    step_s, sample_s, parts, spread = cpu_baseline(g, n, d["m"], rows)
    sel = np.random.default_rng(9).choice(n, rows, replace=False)
    t0 = time.perf_counter()
    box = C.TABLE_RANGES
    O.cool_update(g["rho"][sel], g["u"][sel], Z[sel], J[sel], cfg.dt,
                  (units.unit_density_cgs, units.unit_u_cgs, units.unit_time_s, units.XH, units.u_floor_cgs),
                  [b[0] for b in box], [b[1] for b in box])
    t_cool = time.perf_counter() - t0
    parts["cooling_update"] = t_cool
    step_s += t_cool * n / rows
    return {
        "value": n / step_s, "unit": "particle-steps/s", "cores": threads, "kind": "port",
        "sample": f"{rows} random target rows x all {n} sources of the FP64 all-pairs restatement of "
                  f"sph_serial_v2.py's step (h_smooth_fast, getDensity, getAcc_g_smth, div_curlVel, getAcc_sph_shear, "
                  f"get_dU_shear) plus the exact FP64 cooling update of {rows} particles; "
                  f"{sample_s + t_cool:.1f} s measured, extrapolated x{n / rows:.0f}",
        "step_seconds_extrapolated": step_s, "function_seconds_sample": parts, "sampling": spread,
        "host": host_info(),
    }


def c2_section(args, cfg, run):
    """Evrard collapse N = 1,048,576 without cooling: throughput and roofline, the FP64 form,
    and the first force evaluation against the FP64 oracle's tree path at the device's
    positions and h (neighbour sets compared as integer sets; rho, acceleration, du/dt
    errors), with that oracle path and the all-pairs restatement timed on the host."""
    from sph_3_amd import ic

    n = args.n
    d = ic.evrard(n, seed=42)
    sim, dt = run("mixed", args.steps, args.warmup, d=d)
    st = sim.stats()
    g = sim.get()
    del sim
    ms = 1e3 * dt / args.steps
    roof, extra = roofline(st, n)
    res = {"workload": f"evrard_collapse, n = {n}", "n": n, "ms_per_step": ms, "value": n * args.steps / dt,
           "unit": "particle-steps/s", "roofline": roof, **extra}
    if not args.no_fp64_form:
        steps64 = max(2, args.steps // 2)
        sim64, dt64 = run("fp64", steps64, 1, d=d)
        res["fp64_form"] = {"ms_per_step": 1e3 * dt64 / steps64, "mixed_speedup": (1e3 * dt64 / steps64) / ms}
        del sim64
    log("[bench] 1M: first force evaluation of both forms")
    # the parity evaluation starts from the same particles moving with seeded velocities
    # (0.02 rms per axis), so that viscosity, the Balsara switch and du/dt are exercised
    dv = dict(d, vel=np.random.default_rng(9).normal(size=(n, 3)) * 0.02)
    states = prime_states(dv, cfg, neighbours=not args.no_cpu_baseline)
    res["accel_error_mixed_vs_fp64"] = accel_error(states)
    res["column_density"] = column_section(args, cfg, run, ms, st["ngb_total"] / n, d) if not args.no_column else None
    res["file_io"] = io_section(g, d["m"])
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log(f"[bench] 1M: oracle tree path on {threads} threads")
        tree_s, tree_parts, ref = cpu_tree_oracle(states["mixed"], d["m"], cfg)
        off, idx = states["sets"]
        res["oracle_parity"] = {
            "what": "first force evaluation from the same particles; oracle = FP64 tree neighbours, CSR hydro "
                    "and group-walk gravity (same tree and opening test) at the device's positions and h",
            "neighbour_sets_bit_exact": bool(np.array_equal(off, ref["off"]) and np.array_equal(idx, ref["idx"])),
            "neighbour_entries": int(len(ref["idx"])),
            **oracle_errors(states, ref, dv, cfg.dt),
        }
        res["cpu_tree_oracle"] = {
            "value": n / tree_s, "unit": "particle-steps/s", "cores": threads,
            "kind": "oracle tree path (not the reference's algorithm): tree neighbours, CSR hydro, group-walk "
                    "gravity at the device's h, without the h solve; every row",
            "seconds": tree_s, "function_seconds": tree_parts,
        }
        log(f"[bench] 1M: cpu baseline {args.cpu_rows} rows")
        step_s, sample_s, parts, spread = cpu_baseline(g, n, d["m"], args.cpu_rows)
        res["cpu_baseline"] = {
            "value": n / step_s, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"{args.cpu_rows} random target rows x all {n} sources of the FP64 all-pairs restatement "
                      f"of sph_serial_v2.py's step; {sample_s:.1f} s measured, extrapolated x{n / args.cpu_rows:.0f}",
            "step_seconds_extrapolated": step_s, "function_seconds_sample": parts, "sampling": spread,
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
        roof, _ = roofline(st, n)
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
            "roofline": roof,
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
        roof, _ = roofline(st, n)
        res = {
            "roofline": roof,
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


def column_section(args, cfg, run, ms_step, kbar, d):
    """Column density N_H from |grad rho| (the CLOUDY driver's grad_rho_norm_ngb and N_H
    loop), the per-particle time-step criteria (dt_array_indiv_dt, C_CFL = 0.3) computed in
    every force evaluation of the 1M Evrard step, and the driver's cooling update over
    (n_H, T, r, N_H) fed by that column density (doCloudyHCooling, on a synthetic table of
    similar size), units 1e10 M_sun and 10 kpc: each pass's HIP-event time, the column pass's
    algorithmic bandwidth and the step time with all three against the headline step."""
    from sph_3_amd import column as C

    from sph_3_amd import cooling as K

    units = C.ColumnParams.from_mass_length(1.989e43, 3.0857e22)
    # the reference's cooling over (n_H, T, r, N_H) fed by this column density; its table file
    # is absent, so a smooth stand-in of similar size times the pass
    nh_table = K.NHCoolingTable.synthetic()
    nh_cool = K.CoolingParams.from_mass_length(1.989e43, 3.0857e22, mode="nh")

    def setup(s):
        s.set_column(units)
        s.set_cooling(None, None, nh_cool, surrogate=nh_table)

    cfg.C_CFL = 0.3
    try:
        sim, dt = run("mixed", args.steps, args.warmup, d=d, cooling=setup)
    finally:
        cfg.C_CFL = 0.0
    st = sim.stats()
    n = st["n"]
    ms_col = st["phase_ms"]["column"] / max(1, st["phase_calls"]["column"])
    ms_ts = st["phase_ms"]["timestep"] / max(1, st["phase_calls"]["timestep"])
    ms_nh = st["phase_ms"]["cooling"] / max(1, st["phase_calls"]["cooling"])
    nh_floor = st.get("cool_floor")
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
        "ms_nh_cooling_pass": ms_nh, "nh_cooling_table": f"synthetic {'x'.join(map(str, nh_table.dims))}",
        "nh_cooling_floored_fraction_last_step": nh_floor / n,
        "ms_per_step_with_all": 1e3 * dt / args.steps, "ms_per_step_without": ms_step,
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1048576, help="particles of the Evrard step without cooling")
    ap.add_argument("--cpu-rows", type=int, default=4096)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64-form", action="store_true")
    ap.add_argument("--no-mlp", action="store_true", help="cooling by the table instead of the surrogate")
    ap.add_argument("--mlp-n", type=int, default=4194304)
    ap.add_argument("--mlp-epochs", type=int, default=200, help="training passes of the surrogate")
    ap.add_argument("--cooling-n", type=int, default=4194304, help="particles of the headline step")
    ap.add_argument("--no-c2", action="store_true", help="skip the 1M Evrard block")
    ap.add_argument("--no-c1", action="store_true", help="skip the 10,000-particle block")
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

    from sph_3_amd.cooling import KPC_CM, M_SUN_G
    from sph_3_amd.step import DeviceSPH, StepConfig

    cfg = StepConfig(dt=1e-3, gamma=5 / 3, alpha=1.0, G=1.0, theta=0.5, softening="plummer", eps_plummer=1e-3,
                     n_ngb=64.0, h_tol=1e-10, precision="mixed", balsara=True)

    def run(precision, steps, warmup, d, cooling=None, n_max=None):
        cfg.precision = precision
        sim = DeviceSPH(d["pos"], d["vel"], d["m"], d["u"], d["h0"], cfg, n_max=n_max)
        if cooling is not None:
            cooling(sim)
        log(f"[bench] {precision} n = {len(d['m'])}: prime")
        sim.prime()
        for w in range(warmup):
            sim.step()
        log(f"[bench] {precision}: {warmup} warm-up steps done")
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

    # ---- the cooling surrogate (the headline step's log Lambda), trained on the device
    mlp = None
    if not args.no_mlp:
        log("[bench] cooling surrogate: train, calibrate, time")
        mlp = mlp_section(args.mlp_n, args.mlp_epochs)
    # ---- headline: Evrard collapse + cooling at N = 4,194,304
    d, Z, J, units, cooling = cooling_setup(args)
    n = len(d["m"])
    mode = "mlp" if mlp is not None else "table"
    sim, dt = run("mixed", args.steps, args.warmup, d=d, cooling=cooling(mode))
    st = sim.stats()
    g = sim.get()
    dev_bytes = sim.device_bytes()
    del sim
    ms_step = 1e3 * dt / args.steps
    value = n * args.steps / dt
    roof, extra = roofline(st, n)
    out = {
        "metric": "SPH particle-steps/sec on 1 GPU, mixed-precision vs FP64, at <=1e-3 accel error",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32",
        "precision": "mixed: f32 pair/gravity arithmetic with Kahan f32 sums; f64 positions, velocities, "
                     "centres of mass, h solve, reductions; cooling rate by a bf16 (f32-accumulating) network "
                     "with an f32 table fallback, f64 u update",
        "data": "synthetic (seeded Evrard sphere, seeded per-particle Z and J21)",
        "config": {"workload": "evrard_collapse + cooling", "n": n, "M": 1.0, "R": 1.0, "G": 1.0, "u0": 0.05,
                   "gamma": 5 / 3, "n_ngb": 64, "theta": 0.5, "softening": "plummer", "eps": 1e-3,
                   "alpha": 1.0, "balsara": True, "dt": 1e-3, "cooling": mode, "cooling_tol_dex": units.tol,
                   "units": {"mass_g": 1e10 * M_SUN_G, "length_cm": 10 * KPC_CM}},
        "roofline": roof,
        **extra,
        "cooling_fallback_fraction_last_step": st["cool_fallback"] / n,
        "cooling_floored_fraction_last_step": st["cool_floor"] / n,
        # every buffer the step's context holds (state, tree, rows, h-solve scratch, tables, surrogate)
        "device_bytes": dev_bytes, "device_bytes_per_particle": dev_bytes / n,
    }
    if not args.no_fp64_form:
        steps64 = max(2, args.steps // 2)
        sim64, dt64 = run("fp64", steps64, 1, d=d, cooling=cooling("table"))
        out["fp64_form"] = {"ms_per_step": 1e3 * dt64 / steps64, "mixed_speedup": (1e3 * dt64 / steps64) / ms_step,
                            "cooling_mode": "table (no FP64 network)"}
        del sim64
        log("[bench] headline: acceleration of the mixed form against the FP64 form")
        out["fp64_form"]["accel_error"] = accel_error(prime_states(d, cfg, cooling=cooling("table")))
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cooling_cpu_baseline(args, cfg, g, d, Z, J, units)
        out["speedup_vs_cpu"] = value / out["cpu_baseline"]["value"]
    del g, d
    if mlp is not None:
        out["mlp_cooling"] = mlp
    # ---- the other configurations (each with its own roofline)
    if not args.no_c2:
        out["evrard_1m"] = c2_section(args, cfg, run)
    if not args.no_c1 and not args.no_cpu_baseline:
        log("[bench] the cpu-runnable case, n = 10000")
        out["evrard_10k"] = c1_section(cfg, run)
    if not args.no_agn:
        out["agn_outflow"] = agn_section(args, cfg, run)
    if not args.no_sedov:
        out["sedov_blast"] = sedov_section(args, cfg, run)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
