#!/usr/bin/env python3
"""Benchmark of the TBPM hot path on one MI355X.

Workload (the metric's configuration): pristine graphene 2048 x 2048 with
periodic boundaries (8,388,608 orbitals, 25,165,824 non-zeros), rescaled by
9 eV, 8 random-phase states propagated together for 2048 Chebyshev time
steps (21 Bessel terms at precision 1e-15, i.e. 20 SpMM terms per step) with
the DOS correlation C(t) formed every step.  FP32 storage and arithmetic,
FP64 correlation sums.

A "step" is one time step of all 8 states; ``value`` is step*vectors/s over
the timed steps with the states resident in HBM.  Printed as one JSON line,
with
  * ``roofline``: algorithmic bytes per Chebyshev term
      nnz*(s_val + 4) + 4*(n + 1) + K*5*n*s_vec   (s_vec = 8 B for complex64;
      s_val = 4 B: graphene's hoppings are real and stored as float)
    divided by the average time of one term over the timed region (HIP
    events on the library's stream), against 8 TB/s;
  * ``fp64_k1``: the plain FP64 single-state path on the same sample (the
    speedup denominator) and ``speedup_vs_fp64_k1``;
  * ``ct_err_fp32_vs_fp64``: max |C32(t) - C64(t)| over all timed steps for
    state 0, both from the GPU on the identical initial state;
  * ``cpu_baseline``: the reference's own Fortran tbpm_dos (compiled from its
    sources into oracle/_ref) timed on the host cores on a bounded sample.

``--gpus N`` with N > 1 exits: this build drives one GPU.
"""
import argparse
import json
import math
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
METRIC = "TBPM step·vectors/s + % HBM roofline, 8.4M-orb graphene; C(t) err vs FP64"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2048)
    p.add_argument("--warmup", type=int, default=16)
    p.add_argument("--L", type=int, default=2048)
    p.add_argument("--vectors", type=int, default=8)
    p.add_argument("--fp64-steps", type=int, default=None,
                   help="timed steps of the FP64 K=1 run (default: --steps)")
    p.add_argument("--only-fp32", action="store_true",
                   help="skip the FP64 and CPU legs (for profiling)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--cpu-steps", type=int, default=16)
    p.add_argument("--no-graph", action="store_true")
    p.add_argument("--no-tune", action="store_true",
                   help="use the default launch shape instead of the autotuner")
    p.add_argument("--copy-probe", action="store_true",
                   help="time a plain device copy of known size (counter calibration)")
    p.add_argument("--workload", choices=("dos", "dc", "batch", "ac", "latency"),
                   default="dos",
                   help="dos: the metric's configuration (default); dc: disordered "
                        "field-threaded DC conductivity; batch: 32 uneven instances; "
                        "ac: AC conductivity on the 8.4M-orbital sample; latency: "
                        "131,072-orbital single-state steps")
    p.add_argument("--order", choices=("caller", "tiles"), default="tiles",
                   help="row order of the device layout: the sample's own, or square "
                        "tiles of --tile nm (Sample.tile_order), which also lets the "
                        "tuner choose three or four Chebyshev terms per pass")
    p.add_argument("--tile", type=float, default=5.0,
                   help="dos and batch workloads: tile edge in nm of the spatial row order "
                        "(5.0 nm fits 8 complex64 states a row with three halo rings in "
                        "the LDS of one workgroup)")
    p.add_argument("--cond-tile", type=float, default=None,
                   help="dc and ac workloads: tile edge in nm of a spatial row order "
                        "(0: the sample's own order; default 2.0 for dc, 1.5 for ac)")
    p.add_argument("--cpu-child", action="store_true", help=argparse.SUPPRESS)
    args = p.parse_args()
    if args.cond_tile is None:
        # measured best: DC 16 states 2.0 nm (4328 step·vectors/s, 3749 at
        # 1.5 nm); AC 24 vectors 1.5 nm (535 state·steps/s, 507 at 2.0 nm)
        args.cond_tile = 2.0 if args.workload == "dc" else 1.5
    return args


def copy_probe(nbytes=1 << 30, reps=20):
    """Plain device-to-device copy (torch's copy kernel) of nbytes, timed
    with events: the calibration kernel for FETCH_SIZE / WRITE_SIZE and the
    practical streaming bandwidth of the device."""
    import torch
    x = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda:0").fill_(1.0)
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        y.copy_(x)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    gbs = 2 * nbytes / (ms / 1e3) / 1e9
    print(json.dumps({"copy_bytes": nbytes, "copy_ms": ms, "copy_rw_GBs": gbs}), flush=True)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# CPU baseline: the reference Fortran, in a child process (no GPU there)
# ---------------------------------------------------------------------------
def cpu_child(args):
    # numpy and the oracle only: no torch or HIP runtime (and their OpenMP
    # runtime) in the process that times the reference's OpenMP code
    import numpy as np
    import oracle
    L = args.L
    hi, hj, hv, dr, pos = oracle.graphene_hops_vec(L, L)
    n = len(pos)
    hv = hv * (1.0 / 9.0)
    ip, ix, dt = oracle.build_ham_csr(np.zeros(n), hi, hj, hv)
    bes = oracle.bessel_series(2 * math.pi / 18.0 * 9.0, 1e-15)
    t0 = time.perf_counter()
    oracle.ref_tbpm_dos(bes, ip, ix, dt, 1337, 1, 1, chk_step=0)
    t1 = time.perf_counter()
    oracle.ref_tbpm_dos(bes, ip, ix, dt, 1337, 1 + args.cpu_steps, 1, chk_step=0)
    t2 = time.perf_counter()
    per_step = ((t2 - t1) - (t1 - t0)) / args.cpu_steps
    sys.stdout.flush()
    print(json.dumps({"per_step_s": per_step, "t_short": t1 - t0, "t_long": t2 - t1}),
          flush=True)


def cpu_baseline(args):
    import resource
    if not os.path.exists(os.path.join(HERE, "oracle", "_ref", "libtbpm_ref.so")):
        log("cpu_baseline: oracle/_ref/libtbpm_ref.so missing, skipped")
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
               OMP_PLACES="cores")

    def unlimited_stack():
        # the reference keeps n-sized automatic arrays on the stack
        resource.setrlimit(resource.RLIMIT_STACK,
                           (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

    cmd = [sys.executable, os.path.join(HERE, "bench.py"), "--cpu-child", "--L", str(args.L),
           "--cpu-steps", str(args.cpu_steps)]
    log(f"cpu_baseline: reference tbpm_dos, {threads} threads, {args.cpu_steps} steps")
    r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=900,
                       preexec_fn=unlimited_stack)
    if r.returncode != 0:
        log("cpu_baseline failed:", r.stderr[-2000:])
        return None
    # the Fortran prints progress lines on the same stdout
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith('{"per_step_s"')]
    if not lines:
        log("cpu_baseline: no result line:", r.stdout[-1000:])
        return None
    res = json.loads(lines[-1])
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": 1.0 / res["per_step_s"], "unit": "step·vectors/s", "cores": threads,
            "kind": "reference",
            "sample": (f"reference Fortran tbpm_dos (oracle/_ref, amdflang -O3 -fopenmp) on "
                       f"graphene {args.L}x{args.L}, FP64, 1 state: time of T={1 + args.cpu_steps}"
                       f" minus T=1 over {args.cpu_steps} steps; host {cpu_model}"),
            "per_step_s": res["per_step_s"]}


# ---------------------------------------------------------------------------
# GPU legs
# ---------------------------------------------------------------------------
def layout_name(args, shape):
    """Row order and pass mode of a run (matches the traffic profile)."""
    base = "caller order" if args.order == "caller" else f"tiles of {args.tile} nm"
    return base + passes_name(shape)


def passes_terms(shape):
    """Terms per pass of a tuner shape (1: one term per pass)."""
    return max(1, int(shape[2])) if len(shape) > 2 else 1


def passes_name(shape):
    return {1: ", one term per pass", 3: ", three terms per pass",
            4: ", four terms per pass"}.get(passes_terms(shape), ", several terms per pass")


def shape_json(shape):
    return {"blocks_per_cu": shape[0], "lds_stage": bool(shape[1]),
            "terms_per_pass": passes_terms(shape)}


def term_bytes(n, nnz, K, s, real=False):
    """Algorithmic bytes of one Chebyshev term: H (values + int32 columns +
    row pointers) and five vector streams per state.  s = bytes of one
    complex vector element; H values take s bytes, or s/2 when the
    Hamiltonian is stored with real values (no imaginary parts)."""
    vs = s // 2 if real else s
    return nnz * (vs + 4) + 4 * (n + 1) + K * 5 * n * s


def gpu_run(H, bes, psi, warmup, steps):
    from tbplas_amd import tbpm
    run = tbpm.DosRun(H, bes, psi, warmup + steps)
    if warmup:
        run.advance(warmup)
    ms = run.advance(steps)
    corr = run.result(chk_step=0)
    run.close()
    return ms, corr


def cond_order(s, args):
    """(row order, tile offsets) for the dc and ac workloads: square tiles
    of --cond-tile nm, or the sample's own order."""
    if args.cond_tile > 0:
        return s.tile_order(args.cond_tile, return_tiles=True)
    return None, None


def dc_workload(args):
    """Graphene 1024x1024, 1% vacancies, Anderson W = 1 eV (U[-0.5, 0.5]),
    Peierls phases for B = 20 T (gauge 0), rescale 9; DC correlation at the 5
    grid energies nearest -1, -0.5, 0, 0.5, 1 eV; 16 states, 1024 steps,
    complex64 (FP64 quasi-eigenstates and correlations)."""
    import numpy as np
    import tbplas_amd as tb
    from tbplas_amd import tbpm, _lib, tune
    L, K, T = 1024, 16, args.steps if args.steps != 2048 else 1024
    s = tb.make_graphene_sample(L, vacancies=tb.random_vacancies(2 * L * L, 0.01, 2024))
    tb.add_anderson_disorder(s, 1.0, 2025)
    s.set_magnetic_field(20.0, 0)
    s.rescale_ham(9.0)
    ip, ix, hop, dx, dy = s.build_ham_dxy_device()
    n = s.num_orb
    en = np.array([0.5 * i * s.energy_range / T - s.energy_range / 2. for i in range(2 * T)])
    inds = np.array([int(np.argmin(np.abs(en - e))) for e in (-1.0, -0.5, 0.0, 0.5, 1.0)])
    ts = 2 * math.pi / s.energy_range
    bes = tbpm.bessel_series(ts * s.rescale, 1e-14)
    h = _lib.default_handle(0)
    order, tiles = cond_order(s, args)
    H = tbpm.Hamiltonian(n, ip, ix, hop, dx, dy, cur_scale=s.rescale, precision="fp32", handle=h,
                         order=order, tiles=tiles)
    shape = tune.apply(H, 2 * K, bes, not args.no_tune)
    psi = tb.states.random_states(n, tb.states.sample_seeds(1337, K))
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    d, c = tbpm.corr_dc(H, bes, psi, T, ts, en, inds, chk_thr=1e-4)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # phase 1 propagates 2K states for T steps, phase 2 2K states for T-1 steps
    value = 2 * K * (2 * T - 1) / dt
    out = {"metric": "DC correlation step·vectors/s (disordered field-threaded graphene)",
           "value": round(value, 2), "unit": "step·vectors/s", "n_gpus": 1,
           "higher_is_better": True, "dtype": "fp32", "seconds": round(dt, 3),
           "config": {"workload": f"graphene {L}x{L}, 1% vacancies, Anderson W=1 eV, B=20 T, "
                                  f"DC at 5 energies, {K} states, {T} steps",
                      "n_orbitals": n, "nnz": int(ix.numel()), "energies": en[inds].tolist(),
                      "layout": ("caller order" if order is None
                                 else f"tiles of {args.cond_tile:g} nm")
                      + passes_name(shape),
                      "launch_shape": shape_json(shape)},
           "dc_corr_t1_mean": [float(np.mean(c[:, a, :, 0]).real) for a in range(2)]}
    print(json.dumps(out), flush=True)


def ac_workload(args):
    """Pristine graphene 2048x2048 (rescale 9), AC conductivity correlation
    of 8 states at 300 K, mu = 0: the Fermi and one-minus-Fermi polynomials
    (FFT coefficients at 2^15 points, precision 1e-12) applied once, then
    J_x psi1, J_y psi1 and psi2 of every state (24 vectors) propagated
    together with tau = pi/2 (16 Bessel terms at 1e-14) and the four current
    correlations formed every step; complex64 storage, FP64 sums."""
    import numpy as np
    import torch
    import tbplas_amd as tb
    from tbplas_amd import tbpm, _lib, tune
    L, K = 2048, 8
    T = args.steps
    s = tb.make_graphene_sample(L)
    s.rescale_ham(9.0)
    ip, ix, hop, dx, dy = s.build_ham_dxy_device()
    n = s.num_orb
    ts = 2 * math.pi / s.energy_range * 0.5
    bes = tbpm.bessel_series(ts * s.rescale, 1e-14)
    beta = 1.0 / (8.617333262e-5 * 300.0) * s.rescale
    cF = tbpm.fermi_cheb_coef(2 ** 15, beta, 0.0, False, 1e-12)
    cO = tbpm.fermi_cheb_coef(2 ** 15, beta, 0.0, True, 1e-12)
    h = _lib.default_handle(0)
    h.set_graph(not args.no_graph)
    order, tiles = cond_order(s, args)
    H = tbpm.Hamiltonian(n, ip, ix, hop, dx, dy, cur_scale=s.rescale, precision="fp32",
                         handle=h, order=order, tiles=tiles)
    shape = tune.apply(H, 3 * K, bes, not args.no_tune)
    psi = tb.states.random_states(n, tb.states.sample_seeds(1337, K))
    # Fermi operators alone (T = 1 also forms one correlation step)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tbpm.corr_ac(H, bes, cF, cO, psi, 1, chk_thr=1e-4)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    c = tbpm.corr_ac(H, bes, cF, cO, psi, T, chk_thr=1e-4)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    t_prop = (t2 - t1) - (t1 - t0)
    value = K * (T - 1) / t_prop
    out = {"metric": "AC correlation state·steps/s (8.4M-orbital graphene, 3 propagated "
                     "vectors per state)",
           "value": round(value, 2), "unit": "state·steps/s", "n_gpus": 1,
           "higher_is_better": True, "dtype": "fp32",
           "ms_per_step": round(t_prop / (T - 1) * 1e3, 4),
           "fermi_setup_s": round(t1 - t0, 3), "total_s": round(t2 - t1, 3),
           "config": {"workload": f"graphene {L}x{L} AC, {K} states (24 vectors), {T} steps, "
                                  f"{len(bes)} Bessel terms, Fermi terms {len(cF)}",
                      "n_orbitals": n, "nnz": int(ix.numel()),
                      "layout": ("caller order" if order is None
                                 else f"tiles of {args.cond_tile:g} nm")
                      + passes_name(shape),
                      "launch_shape": shape_json(shape)},
           "corr_xx_t2_mean": complex(np.mean(c[:, 0, 1])).real}
    H.close()
    print(json.dumps(out), flush=True)


def latency_workload(args):
    """Pristine graphene 256x256 (131,072 orbitals, rescale 9), one state,
    1024 DOS steps (21 Bessel terms at 1e-15): the case whose vectors fit in
    the last-level cache, reported as microseconds per step, complex128 and
    complex64, hipGraph replay, 2 nm tiles (the tuner picks one or three
    terms per pass)."""
    import tbplas_amd as tb
    from tbplas_amd import tbpm, _lib, tune
    L, T = 256, 1024
    s = tb.make_graphene_sample(L)
    s.rescale_ham(9.0)
    ip, ix, dt = s.build_ham_csr_device()
    n = s.num_orb
    bes = tbpm.bessel_series(2 * math.pi / s.energy_range * s.rescale, 1e-15)
    psi = tb.states.random_states(n, tb.states.sample_seeds(1337, 1))
    h = _lib.default_handle(0)
    h.set_graph(not args.no_graph)
    # 2 nm tiles (~150 rows): enough workgroups for 256 CUs on this size
    order, tiles = (s.tile_order(2.0, return_tiles=True) if args.order == "tiles"
                    else (None, None))
    res = {}
    for prec in ("fp64", "fp32"):
        H = tbpm.Hamiltonian(n, ip, ix, dt, precision=prec, handle=h, order=order, tiles=tiles)
        shape = tune.apply(H, 1, bes, not args.no_tune)
        ms, _ = gpu_run(H, bes, psi, 32, T)
        H.close()
        res[prec] = {"us_per_step": round(ms / T * 1e3, 2),
                     "us_per_term": round(ms / T / (len(bes) - 1) * 1e3, 3),
                     "layout": ("caller order" if order is None else "tiles of 2.0 nm")
                     + passes_name(shape),
                     "launch_shape": shape_json(shape)}
    out = {"metric": "TBPM microseconds per step, 131,072-orbital graphene, 1 state",
           "value": res["fp64"]["us_per_step"], "unit": "us/step", "n_gpus": 1,
           "higher_is_better": False, "dtype": "fp64",
           "config": {"workload": f"graphene {L}x{L} DOS, 1 state, {T} steps",
                      "n_orbitals": n, "nnz": int(ix.numel()), "hipgraph": not args.no_graph},
           "fp64": res["fp64"], "fp32": res["fp32"]}
    print(json.dumps(out), flush=True)


def batch_workload(args):
    """32 independent disordered graphene instances of uneven size (edges from
    {128, 256, 384, 512, 768, 1024}, vacancies 0-3%, Anderson W 0-2 eV, all
    seeded), rescaled per instance by the row-sum bound, 4 states each, DOS
    correlation for 1024 steps, every instance in one launch per pass (rows
    in --tile nm tiles: three Chebyshev terms per pass where the tuner finds
    that faster), replayed from a hipGraph; complex64, with the
    per-instance error against the complex128 run on the same states."""
    import numpy as np
    import tbplas_amd as tb
    from tbplas_amd import tbpm, _lib, tune
    rng = np.random.default_rng(4)
    edges = rng.choice([128, 256, 384, 512, 768, 1024], 32)
    vac = rng.uniform(0.0, 0.03, 32)
    Ws = rng.uniform(0.0, 2.0, 32)
    K, T = 4, args.steps if args.steps != 2048 else 1024
    h = _lib.default_handle(0)
    insts = []
    for i in range(32):
        L = int(edges[i])
        s = tb.make_graphene_sample(L, vacancies=tb.random_vacancies(2 * L * L, vac[i], 100 + i))
        tb.add_anderson_disorder(s, Ws[i], 200 + i)
        s.rescale_ham()
        ip, ix, dt = s.build_ham_csr_device()
        insts.append((s.num_orb, int(ix.numel()), ip, ix, dt,
                      tb.states.random_states(s.num_orb, tb.states.sample_seeds(300 + i, K)),
                      s.tile_order(args.tile, return_tiles=True) if args.order == "tiles"
                      else (None, None)))
    bes = tbpm.bessel_series(math.pi, 1e-14)
    res = {}
    for prec in ("fp32", "fp64"):
        hams = [tbpm.Hamiltonian(n, ip, ix, dt, precision=prec, handle=h, order=ot[0],
                                 tiles=ot[1])
                for (n, nnz, ip, ix, dt, _, ot) in insts]
        shape = tune.apply(hams, K, bes, not args.no_tune)
        run = tbpm.DosRun(hams, bes, [x[5] for x in insts], T + 16)
        run.advance(16)
        ms = run.advance(T)
        res[prec] = (ms, run.result(chk_step=0), shape,
                     all(H.info()["real_values"] for H in hams))
        run.close()
        for H in hams:
            H.close()
    ms, c32, shape, _ = res["fp32"]
    c64 = res["fp64"][1]
    errs = [float(np.max(np.abs(a - b))) for a, b in zip(c32, c64)]
    n_tot = sum(x[0] for x in insts)
    nnz_tot = sum(x[1] for x in insts)
    t_step = ms / 1e3 / T
    nterm = len(bes) - 1
    bt = (term_bytes(n_tot, nnz_tot, K, 8, res["fp32"][3]) - 4 * (n_tot + 1)
          + sum(4 * (x[0] + 1) for x in insts))
    ach = bt / (t_step / nterm) / 1e9
    out = {"metric": "uneven batch DOS step·vectors/s (32 instances in one launch per pass)",
           "value": round(32 * K / t_step, 2), "unit": "instance-step·vectors/s", "n_gpus": 1,
           "higher_is_better": True, "dtype": "fp32", "ms_per_step": round(t_step * 1e3, 4),
           "config": {"workload": f"32 disordered graphene instances, edges {sorted(set(edges.tolist()))}, "
                                  f"{K} states each, {T} steps", "n_orbitals_total": n_tot,
                      "nnz_total": nnz_tot, "launch_shape": shape_json(shape),
                      "layout": layout_name(args, shape)},
           "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4)},
           "fp64_ms_per_step": round(res["fp64"][0] / T, 4),
           "ct_err_fp32_vs_fp64_per_instance_max": max(errs),
           "ct_err_fp32_vs_fp64_per_instance_median": float(np.median(errs))}
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.cpu_child:
        cpu_child(args)
        return 0
    if args.copy_probe:
        copy_probe()
        return 0
    if args.workload != "dos":
        if args.gpus != 1:
            print("bench.py drives one GPU", file=sys.stderr)
            return 2
        {"dc": dc_workload, "batch": batch_workload, "ac": ac_workload,
         "latency": latency_workload}[args.workload](args)
        return 0
    if args.gpus != 1:
        print(f"bench.py drives one GPU; --gpus {args.gpus} is not supported "
              "(states are independent: several GPUs would each propagate a share "
              "of them and sum the correlation partials once per run).", file=sys.stderr)
        return 2
    cpu = None if (args.no_cpu or args.only_fp32) else cpu_baseline(args)

    import numpy as np
    import tbplas_amd as tb
    from tbplas_amd import tbpm, _lib

    t0 = time.perf_counter()
    s = tb.make_graphene_sample(args.L)
    s.rescale_ham(9.0)
    indptr, indices, data = s.build_ham_csr_device()
    n, nnz = s.num_orb, int(indices.numel())
    ts = 2 * math.pi / s.energy_range
    bes = tbpm.bessel_series(ts * s.rescale, 1e-15)
    K = args.vectors
    psi = tb.states.random_states(n, tb.states.sample_seeds(1337, K))
    log(f"setup {time.perf_counter() - t0:.1f}s: n={n} nnz={nnz} n_bes={len(bes)}")
    h = _lib.default_handle(0)
    h.set_graph(not args.no_graph)

    from tbplas_amd import tune
    order, tiles = (s.tile_order(args.tile, return_tiles=True) if args.order == "tiles"
                    else (None, None))
    H32 = tbpm.Hamiltonian(n, indptr, indices, data, precision="fp32", handle=h, order=order,
                           tiles=tiles)
    real = H32.info()["real_values"]
    shape32 = tune.apply(H32, K, bes, not args.no_tune)
    ms32, c32 = gpu_run(H32, bes, psi, args.warmup, args.steps)
    H32.close()
    t_step = ms32 / 1e3 / args.steps
    value = K / t_step
    nterm = len(bes) - 1
    bt = term_bytes(n, nnz, K, 8, real)
    achieved = bt / (t_step / nterm) / 1e9
    log(f"fp32 K={K}: {ms32:.1f} ms for {args.steps} steps -> {value:.1f} step·vectors/s, "
        f"{achieved:.0f} GB/s algorithmic")

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "step·vectors/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t_step * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded random-phase states; pristine graphene lattice)",
        "config": {"workload": (f"graphene {args.L}x{args.L} PBC DOS correlation, {K} states "
                                f"together, {args.warmup + args.steps} Chebyshev steps, "
                                "complex64 storage/FP64 correlation"),
                   "n_orbitals": n, "nnz": nnz, "states": K, "bessel_terms": len(bes),
                   "terms_per_step": nterm, "hipgraph": not args.no_graph,
                   "layout": layout_name(args, shape32),
                   "launch_shape": shape_json(shape32)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "bytes_per_term": bt,
                     "term_avg_us": round(t_step / nterm * 1e6, 2)},
    }
    layout = layout_name(args, shape32)
    if passes_terms(shape32) > 1:
        out["roofline"]["note"] = (
            "achieved = the one-term-per-pass algorithmic bytes per term over the "
            "measured time per term; several terms per pass read the state vectors "
            "fewer times, so it can exceed what one term per pass could reach "
            "(traffic: bytes per term actually moved, from the counters)")
    traffic_file = os.path.join(HERE, "profiles", "traffic_term_fp32_k8.json")
    if os.path.exists(traffic_file) and args.L == 2048 and K == 8:
        with open(traffic_file) as f:
            tr = json.load(f)
        if tr.get("layout") == layout and tr.get("bytes_per_term"):
            moved = float(tr["bytes_per_term"])
            out["roofline"]["traffic"] = moved
            # the bytes the counters saw per term over the same time per term
            rate = moved / (t_step / nterm) / 1e9
            out["roofline"]["moved_GBs"] = round(rate, 1)
            out["roofline"]["moved_frac"] = round(rate / HBM_PEAK_GBS, 4)

    if not args.only_fp32:
        steps64 = args.fp64_steps or args.steps
        H64 = tbpm.Hamiltonian(n, indptr, indices, data, precision="fp64", handle=h,
                               order=order, tiles=tiles)
        shape64 = tune.apply(H64, 1, bes, not args.no_tune)
        ms64, c64 = gpu_run(H64, bes, psi[:1], args.warmup, steps64)
        real64 = H64.info()["real_values"]
        H64.close()
        t64 = ms64 / 1e3 / steps64
        b64 = term_bytes(n, nnz, 1, 16, real64)
        tmax = min(args.warmup + args.steps, args.warmup + steps64)
        err = float(np.max(np.abs(c32[0, :tmax + 1] - c64[0, :tmax + 1])))
        out["fp64_k1"] = {"value": round(1.0 / t64, 3), "unit": "step·vectors/s",
                          "ms_per_step": round(t64 * 1e3, 4),
                          "roofline_frac": round(b64 / (t64 / nterm) / 1e9 / HBM_PEAK_GBS, 4),
                          "launch_shape": shape_json(shape64)}
        out["speedup_vs_fp64_k1"] = round(value * t64, 3)
        out["ct_err_fp32_vs_fp64"] = err
        out["ct_err_steps"] = tmax
        log(f"fp64 K=1: {1.0 / t64:.1f} step/s; fp32 vs fp64 max |dC| = {err:.3e}")
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
