#!/usr/bin/env python3
"""Benchmark of the TBPM hot path on one MI355X.

Headline workload (the metric's configuration): pristine graphene 2048 x
2048 with periodic boundaries (8,388,608 orbitals, 25,165,824 non-zeros),
rescaled by 9 eV, 8 random-phase states propagated together (21 Bessel
terms at precision 1e-15, i.e. 20 SpMM terms per time step) with the DOS
correlation C(t) formed every step; rows in square spatial tiles (the
library's tile rule for the leg's row width and terms per pass: 4.8 nm
for 8 complex64 states with five terms) so the tuner can run several
Chebyshev terms per pass.  complex64
storage and arithmetic, FP64 correlation sums.

A "step" is one time step of all 8 states; ``value`` is step*vectors/s over
the timed steps (--steps, after --warmup untimed ones) with the states
resident in HBM, replayed from a hipGraph.  The one JSON line also holds

  * ``roofline``: the dominant kernel (largest share of the step's device
    time) from a segment of --profile-steps steps run right after the
    timed ones with HIP events around every launch on the library's stream:
    its average launch duration and the bytes one launch must move (for the
    tile pass: the workgroup tables, the distinct rows of v_{m-1} its tiles
    gather, v_{m-2} on the tiles and their halo rings, and the output read
    plus written and the last two terms written on the owned rows, counted
    from the tables by the library), ``frac`` = those bytes per second over
    8 TB/s.  ``traffic`` is the counter-measured L2-to-HBM bytes per launch
    (FETCH_SIZE and WRITE_SIZE, copy-calibrated) from profiles/traffic.json
    when that file was measured on the sources the workload's kernels are
    built from and on the same tiles, else null.  ``one_term_equivalent`` restates the speed as one-term-per-pass
    algorithmic bytes (five vector streams per state and H once per term)
    per measured time per term;
  * ``fp64_k8`` (the same 8 states in complex128, 3.4 nm tiles) and
    ``fp64_k1`` (one complex128 state, 7.5 nm tiles: the speedup
    denominator);
  * ``ct_err_fp32_vs_exact`` / ``ct_err_fp64_vs_exact``: max |C(t) - C_exact|
    over the 8 states at sampled steps of a 2048-step run (the timed steps
    are part of it), C_exact from k-space propagation of the same states
    (exact for pristine periodic graphene), computed on the host;
  * ``cpu_baseline``: the reference's own Fortran tbpm_dos (compiled from its
    sources into oracle/_ref) on the host's physical cores;
  * sub-records for the other workloads: ``ac`` (AC conductivity on the same
    sample, 4 states per run), ``dc`` (disordered field-threaded DC),
    ``batch`` (32 uneven instances), ``latency`` (131,072 orbitals),
    ``assembly`` (GPU CSR assembly at 8.4M orbitals and 10^7 atoms) and
    ``hall`` (the Kubo-Bastin moments of the reference's Hall example, on
    the FP64 matrix cores), each with its dominant kernel's roofline.

``--workload X`` runs one of them alone (for profiling); ``--gpus N`` with
N > 1 exits: this build drives one GPU.
"""
import argparse
import hashlib
import json
import math
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6      # dense FP64 matrix-core peak of one MI355X
METRIC = "TBPM step·vectors/s + % HBM roofline, 8.4M-orb graphene; C(t) err vs FP64"
# steps of the accuracy run and the steps compared with the exact propagation
ACC_STEPS = 2048
EXACT_PTS = sorted(set([0, 1, 2, 3, 4, 5, 17, 21, 100, 511, 1023, 1025, 2047, ACC_STEPS]
                       + list(range(0, ACC_STEPS + 1, 128))))
# the AC accuracy run and its sampled steps (t >= 2; column t = 1 is zero)
AC_ACC_STEPS = 256
AC_PTS = [2, 3, 4, 17, 64, 128, 255, AC_ACC_STEPS]
WORKLOADS = ("all", "dos", "ac", "dc", "batch", "latency", "assembly", "hall")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=512)
    p.add_argument("--warmup", type=int, default=16)
    p.add_argument("--profile-steps", type=int, default=16,
                   help="steps after the timed ones with per-launch events (roofline)")
    p.add_argument("--accuracy-steps", type=int, default=ACC_STEPS,
                   help="the DOS runs continue untimed to this many steps for the "
                        "error against the exact propagation (0: no error check)")
    p.add_argument("--L", type=int, default=2048)
    p.add_argument("--vectors", type=int, default=8)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--cpu-steps", type=int, default=16)
    p.add_argument("--no-graph", action="store_true")
    p.add_argument("--no-tune", action="store_true",
                   help="use the default launch shape instead of the autotuner")
    p.add_argument("--copy-probe", action="store_true",
                   help="time a plain device copy of known size (counter calibration)")
    p.add_argument("--workload", choices=WORKLOADS, default="all")
    p.add_argument("--order", choices=("caller", "tiles"), default="tiles",
                   help="row order of the device layout: the sample's own, or square "
                        "tiles (Sample.tile_order), which lets the tuner choose three, "
                        "four or five Chebyshev terms per pass")
    p.add_argument("--tile", type=float, default=None,
                   help="dos and batch: tile edge in nm (default: the library's "
                        "auto_tile_nm for each leg's states and precision, as the Solver's "
                        "DOS driver picks it, to 0.1 nm: 4.8 nm for 8 complex64 states a "
                        "row, 3.2 nm for 8 complex128 states, 7.5 nm for one, 7.1 nm for "
                        "the batch's 4 complex64 states)")
    p.add_argument("--cond-tile", type=float, default=None,
                   help="dc and ac: tile edge in nm (0: the sample's own order; default: "
                        "the library's rule for the run's vectors, as the Solver's "
                        "gpu['dc_tile_nm'] and gpu['ac_tile_nm'] 'auto': dc 3.2 nm for 8 "
                        "states, ac 3.8 nm for 4 complex64 states, 2.5 for complex128)")
    p.add_argument("--dc-states", type=int, default=8,
                   help="dc: states per run (two propagated vectors each), as the "
                        "Solver's gpu['dc_states']")
    p.add_argument("--ac-states", type=int, default=4,
                   help="ac: states per run (three propagated vectors each), as the "
                        "Solver's gpu['ac_states']")
    p.add_argument("--cpu-child", action="store_true", help=argparse.SUPPRESS)
    p.add_argument("--exact-child", action="store_true", help=argparse.SUPPRESS)
    return p.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


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


# ---------------------------------------------------------------------------
# Host side: the reference Fortran's time and the exact correlations, each in
# a child process that loads neither torch nor the HIP runtime
# ---------------------------------------------------------------------------
def physical_cores():
    """(logical CPUs, physical cores) this process may run on."""
    cpus = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in cpus:
        base = f"/sys/devices/system/cpu/cpu{c}/topology/"
        try:
            with open(base + "physical_package_id") as f:
                pkg = f.read().strip()
            with open(base + "core_id") as f:
                core = f.read().strip()
            cores.add((pkg, core))
        except OSError:
            cores.add(("cpu", c))
    return len(cpus), len(cores)


def cpu_model():
    name = ""
    try:
        r = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20)
        for line in r.stdout.splitlines():
            if line.startswith("Model name:"):
                name = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    if not name:
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        name = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    return name or "unknown"


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


def exact_child(args):
    """DOS and AC correlations of the headline states by exact k-space
    propagation (pristine periodic graphene is block diagonal in k; the AC
    Fermi operators are the same Chebyshev series, summed on the band
    energies)."""
    import numpy as np
    import oracle
    from tbplas_amd import states
    L, K = args.L, args.vectors
    psi = states.random_states(2 * L * L, states.sample_seeds(1337, K))
    hi, hj, hv, dr, pos = oracle.graphene_hops_vec(L, L)
    hv = hv * (1.0 / 9.0)
    c = oracle.kspace_graphene_corr_multi(L, L, hv[0], psi, math.pi, EXACT_PTS)
    beta = 1.0 / (8.617333262e-5 * 300.0) * 9.0
    cF = oracle.orc_fermi_coef(2 ** 15, beta, 0.0, False, 1e-12)
    cO = oracle.orc_fermi_coef(2 ** 15, beta, 0.0, True, 1e-12)
    a = oracle.kspace_graphene_ac(L, L, hv[:3], dr[:3], 9.0, cF, cO, psi, math.pi / 2, AC_PTS)
    print(json.dumps({"pts": EXACT_PTS, "re": np.real(c).tolist(), "im": np.imag(c).tolist(),
                      "ac_pts": AC_PTS, "ac_re": np.real(a).tolist(),
                      "ac_im": np.imag(a).tolist()}), flush=True)


def cpu_baseline(args):
    import resource
    if not os.path.exists(os.path.join(HERE, "oracle", "_ref", "libtbpm_ref.so")):
        log("cpu_baseline: oracle/_ref/libtbpm_ref.so missing, skipped")
        return None
    logical, phys = physical_cores()
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    # every physical core of the affinity mask, within the CPU share the
    # machine assigns to the process (OMP_NUM_THREADS) when it sets one
    threads = min(phys, share) if share > 0 else phys
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
               OMP_PLACES="cores")

    def unlimited_stack():
        # the reference keeps n-sized automatic arrays on the stack
        resource.setrlimit(resource.RLIMIT_STACK,
                           (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

    cmd = [sys.executable, os.path.join(HERE, "bench.py"), "--cpu-child", "--L", str(args.L),
           "--cpu-steps", str(args.cpu_steps)]
    log(f"cpu_baseline: reference tbpm_dos, {threads} threads ({phys} physical cores, "
        f"{logical} logical CPUs available), {args.cpu_steps} steps")
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
    model = cpu_model()
    return {"value": 1.0 / res["per_step_s"], "unit": "step·vectors/s", "cores": threads,
            "kind": "reference",
            "sample": (f"reference Fortran tbpm_dos (oracle/_ref, amdflang -O3 -fopenmp) on "
                       f"graphene {args.L}x{args.L}, FP64, 1 state: time of T={1 + args.cpu_steps}"
                       f" minus T=1 over {args.cpu_steps} steps; {threads} OpenMP threads on "
                       f"{model}"),
            "physical_cores_available": phys, "logical_cpus_available": logical,
            "cpu_share": share or None, "cpu_model": model,
            "per_step_s": res["per_step_s"]}


def start_exact(args):
    env = dict(os.environ, OMP_NUM_THREADS="4")
    cmd = [sys.executable, os.path.join(HERE, "bench.py"), "--exact-child", "--L", str(args.L),
           "--vectors", str(args.vectors)]
    return subprocess.Popen(cmd, env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                            text=True)


def finish_exact(proc):
    """{"dos": (steps, (K, steps) exact C(t)), "ac": (steps, (K, 4, steps))}."""
    import numpy as np
    out, err = proc.communicate(timeout=900)
    lines = [ln for ln in out.splitlines() if ln.startswith('{"pts"')]
    if proc.returncode != 0 or not lines:
        log("exact k-space correlations failed:", err[-2000:])
        return None
    d = json.loads(lines[-1])
    return {"dos": (d["pts"], np.array(d["re"]) + 1j * np.array(d["im"])),
            "ac": (d["ac_pts"], np.array(d["ac_re"]) + 1j * np.array(d["ac_im"]))}


# ---------------------------------------------------------------------------
# GPU helpers
# ---------------------------------------------------------------------------
# Sources a workload's kernels are not built from: a change there leaves
# its counter profile valid (every other library source and the header
# count).  The Hall moments run the one-term kernel and k_gram, no tile
# pass; the propagation workloads run no Gram block.
_NOT_BUILT_FROM = {
    "hall": ("tbg_mlp.hip", "tbg_super.hip", "tbg_static.hip", "tbg_block_pass.h",
             "tbg_block.hip", "tbg_pass_"),
    "prop": ("tbg_mlp.hip", "tbg_super.hip", "tbg_static.hip", "tbg_hall.hip"),
}


def src_hash(workload=None):
    """Hash of the library sources a workload's kernels are built from
    (every source when workload is None) and the header: a counter profile
    applies to the build it was measured on."""
    skip = () if workload is None else _NOT_BUILT_FROM[
        "hall" if workload.startswith("hall") else "prop"]
    h = hashlib.sha256()
    csrc = os.path.join(HERE, "tbplas_amd", "csrc")
    for name in sorted(os.listdir(csrc)):
        if any(name.startswith(x) for x in skip):
            continue
        with open(os.path.join(csrc, name), "rb") as f:
            h.update(name.encode() + b"\0" + f.read())
    with open(os.path.join(HERE, "include", "tbplas_amd.h"), "rb") as f:
        h.update(f.read())
    return h.hexdigest()[:16]


def counter_traffic(workload, kernel):
    """Counter-measured bytes per launch of `kernel` for `workload` from
    profiles/traffic.json, when measured on these kernel sources."""
    path = os.path.join(HERE, "profiles", "traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    e = d.get(workload, {}).get(kernel)
    if not e or e.get("src_hash") != src_hash(workload):
        return None
    return e


def passes_terms(shape):
    """Terms per pass of a tuner shape (1: one term per pass)."""
    return max(1, int(shape[2])) if len(shape) > 2 else 1


def passes_name(shape):
    return {1: ", one term per pass", 3: ", three terms per pass", 4: ", four terms per pass",
            5: ", five terms per pass"}.get(passes_terms(shape), ", several terms per pass")


def shape_json(shape):
    return {"blocks_per_cu": shape[0], "lds_stage": bool(shape[1]),
            "terms_per_pass": passes_terms(shape)}


def term_bytes(n, nnz, K, s, real=False):
    """One-term-per-pass algorithmic bytes of one Chebyshev term: H (values +
    int32 columns + row pointers) and five vector streams per state.  s =
    bytes of one complex vector element; H values take s bytes, or s/2 when
    the Hamiltonian is stored with real values."""
    vs = s // 2 if real else s
    return nnz * (vs + 4) + 4 * (n + 1) + K * 5 * n * s


def term_kernel_bytes(n, nnz, K, s, real=False):
    """What the one-term kernel must move per term: H, the gathered v_{m-1},
    v_{m-2} read and v_m written, and the step output's read-modify-write
    every third term (two streams per three terms)."""
    vs = s // 2 if real else s
    return nnz * (vs + 4) + 4 * (n + 1) + K * n * s * (3.0 + 2.0 / 3.0)


def kernel_label(kind, prec, K, real, threads):
    """Template name of a timed launch kind (k_block_pass<R, V, RV, mode,
    threads, terms>)."""
    if kind.startswith("pass"):
        mode, terms = kind[4:].split("x")
        slots = 1
        while slots < K:
            slots <<= 1
        p3 = 3
        while p3 < K:
            p3 <<= 1
        slots = min(slots, p3)
        V = 3 if slots % 3 == 0 else (slots if slots < 2 else 2)
        R = "float" if prec == "fp32" else "double"
        # the launch narrows 16-wave workgroups where the instance needs more
        # registers than a lane has at 16 waves (complex128 states with the
        # dot epilogue after three or more terms, three complex128 states per
        # lane)
        wide = V == 1 or (V == 3 and prec == "fp32") or (
            V == 2 and (prec == "fp32" or (real and not (mode == "4" and int(terms) >= 3))))
        nt = threads if (wide or threads != 1024) else 512
        return (f"k_block_pass<{R}, {V}, {str(bool(real)).lower()}, {mode}, {nt}, "
                f"{terms}>")
    return {"term": "k_term", "ac_corr": "k_accorr", "dc_qe": "k_qe",
            "dc_dot": "k_dcdot"}.get(kind, kind)


def roofline(kt, bytes_of, label_of, workload, extra=None):
    """Roofline record of the dominant timed kernel kind: bytes_of(kind) =
    bytes one launch must move (None: not modelled), label_of(kind) =
    kernel name."""
    tot = sum(v[0] for v in kt.values())
    kind = max(kt, key=lambda k: kt[k][0])
    ms, cnt = kt[kind]
    avg_s = ms / cnt / 1e3
    b = bytes_of(kind)
    name = label_of(kind)
    rec = {"bound": "hbm", "kernel": name, "launches": cnt,
           "launch_avg_us": round(avg_s * 1e6, 2),
           "share_of_device_time": round(ms / tot, 4) if tot > 0 else None,
           "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
           "traffic": None, "bytes_per_launch": None, "traffic_key": workload}
    if b is not None:
        ach = b / avg_s / 1e9
        rec.update({"achieved": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4),
                    "bytes_per_launch": int(b)})
    tr = counter_traffic(workload, name)
    if tr is not None:
        rec["traffic"] = tr["bytes_per_launch"]
        rate = tr["bytes_per_launch"] / avg_s / 1e9
        rec["traffic_GBs"] = round(rate, 1)
        rec["traffic_frac"] = round(rate / HBM_PEAK_GBS, 4)
        rec["traffic_source"] = tr.get("summary")
    else:
        rec["traffic_note"] = ("no counter profile of this build of the kernel sources in "
                               "profiles/traffic.json")
    rec["launches_by_kind"] = {label_of(k): {"launches": v[1],
                                             "avg_us": round(v[0] / v[1] * 1e3, 2)}
                               for k, v in sorted(kt.items(), key=lambda kv: -kv[1][0])}
    if extra:
        rec.update(extra)
    return rec


def timed_run(run, h, args, total):
    """warmup, timed steps (graph replay), profile steps (per-launch events),
    then untimed to `total` steps; returns (timed ms, kernel times)."""
    if args.warmup:
        run.advance(args.warmup)
    ms = run.advance(args.steps)
    kt = {}
    if args.profile_steps > 0:
        h.set_kernel_timing(True)
        try:
            run.advance(args.profile_steps)
            kt = h.kernel_times()
        finally:
            h.set_kernel_timing(False)
    rest = total - run.done
    if rest > 0:
        run.advance(rest)
    return ms, kt


def pass_bytes_fn(sched, n, nnz, K, s, real):
    """bytes_of for propagation launches: a full tile pass from the
    library's table counts, the one-term kernel from its streams."""
    full = None if sched is None else sched["pass_bytes"]

    def f(kind):
        if kind == "term":
            return term_kernel_bytes(n, nnz, K, s, real)
        if kind.startswith("pass0x"):
            return full
        return None
    return f


# ---------------------------------------------------------------------------
# Headline: DOS on 8.4M-orbital graphene
# ---------------------------------------------------------------------------
def leg_tile(args, K, prec):
    """Tile edge (nm) of a DOS leg: --tile, else the library's rule (to
    0.1 nm)."""
    from tbplas_amd.sample import auto_tile_nm
    return args.tile if args.tile is not None else round(auto_tile_nm(K, prec), 1)


def traffic_key(workload, tile):
    """Key of a workload's counter profile: the layout changes the bytes."""
    return f"{workload}@{tile:g}nm" if tile else f"{workload}@caller"


def dos_leg(args, s, csr, tile, bes, psi, prec, K, h, total):
    from tbplas_amd import tbpm, tune
    ip, ix, dt = csr
    ot = s.tile_order(tile, return_tiles=True) if tile else (None, None)
    n, nnz = s.num_orb, int(ix.numel())
    H = tbpm.Hamiltonian(n, ip, ix, dt, precision=prec, handle=h, order=ot[0], tiles=ot[1])
    real = H.info()["real_values"]
    shape = tune.apply(H, K, bes, not args.no_tune)
    threads = H.blocking_launch(K)[1]
    run = tbpm.DosRun(H, bes, psi[:K], total)
    ms, kt = timed_run(run, h, args, total)
    corr = run.result(chk_step=0)
    sched = run.blocking_bytes()
    run.close()
    H.close()
    sz = 8 if prec == "fp32" else 16
    t_step = ms / 1e3 / args.steps
    nterm = len(bes) - 1
    label = lambda k: kernel_label(k, prec, K, real, threads)  # noqa: E731
    rl = roofline(kt, pass_bytes_fn(sched, n, nnz, K, sz, real), label,
                  traffic_key(f"dos_{prec}_k{K}", tile)) if kt else None
    one = term_bytes(n, nnz, K, sz, real) / (t_step / nterm) / 1e9
    if rl is not None:
        rl["one_term_equivalent"] = {
            "GBs": round(one, 1), "frac": round(one / HBM_PEAK_GBS, 4),
            "bytes_per_term": term_bytes(n, nnz, K, sz, real),
            "term_avg_us": round(t_step / nterm * 1e6, 2),
            "note": "one-term-per-pass algorithmic bytes (five vector streams per state, "
                    "H once) per measured time per term; several terms per pass move fewer "
                    "bytes, so this may exceed the device's bandwidth"}
        if sched is not None:
            rl["schedule"] = sched
    return {"value": K / t_step, "ms_per_step": t_step * 1e3, "shape": shape, "corr": corr,
            "roofline": rl, "real": real,
            "layout": (f"tiles of {tile:g} nm" if tile else "caller order") + passes_name(shape)}


def headline(args, cpu, exact_proc):
    import numpy as np
    import tbplas_amd as tb
    from tbplas_amd import tbpm, _lib

    t0 = time.perf_counter()
    s = tb.make_graphene_sample(args.L)
    s.rescale_ham(9.0)
    csr = s.build_ham_csr_device()
    n, nnz = s.num_orb, int(csr[1].numel())
    bes = tbpm.bessel_series(2 * math.pi / s.energy_range * s.rescale, 1e-15)
    K = args.vectors
    psi = tb.states.random_states(n, tb.states.sample_seeds(1337, K))
    log(f"setup {time.perf_counter() - t0:.1f}s: n={n} nnz={nnz} n_bes={len(bes)}")
    h = _lib.default_handle(0)
    h.set_graph(not args.no_graph)
    total = max(args.accuracy_steps, args.warmup + args.steps + args.profile_steps)

    tiles = args.order == "tiles"
    r32 = dos_leg(args, s, csr, leg_tile(args, K, "fp32") if tiles else 0, bes, psi, "fp32", K,
                  h, total)
    log(f"fp32 K={K}: {r32['value']:.1f} step·vectors/s ({r32['ms_per_step']:.3f} ms/step)")
    r64 = dos_leg(args, s, csr, leg_tile(args, K, "fp64") if tiles else 0, bes, psi, "fp64", K,
                  h, total)
    log(f"fp64 K={K}: {r64['value']:.1f} step·vectors/s ({r64['ms_per_step']:.3f} ms/step)")
    a1 = argparse.Namespace(**vars(args))
    a1.profile_steps = 0
    r1 = dos_leg(a1, s, csr, leg_tile(args, 1, "fp64") if tiles else 0, bes, psi, "fp64", 1,
                 h, args.warmup + args.steps)
    log(f"fp64 K=1: {r1['value']:.1f} steps/s")

    out = {
        "metric": METRIC, "value": round(r32["value"], 3), "unit": "step·vectors/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(r32["ms_per_step"], 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded random-phase states; pristine graphene lattice)",
        "config": {"workload": (f"graphene {args.L}x{args.L} PBC DOS correlation, {K} states "
                                f"together, complex64 storage/FP64 correlation"),
                   "n_orbitals": n, "nnz": nnz, "states": K, "bessel_terms": len(bes),
                   "terms_per_step": len(bes) - 1, "hipgraph": not args.no_graph,
                   "layout": r32["layout"],
                   "launch_shape": shape_json(r32["shape"]),
                   "run_steps": total},
        "roofline": r32["roofline"],
        "fp64_k8": {"value": round(r64["value"], 3), "unit": "step·vectors/s",
                    "ms_per_step": round(r64["ms_per_step"], 4),
                    "layout": r64["layout"],
                    "launch_shape": shape_json(r64["shape"]), "roofline": r64["roofline"]},
        "fp64_k1": {"value": round(r1["value"], 3), "unit": "step·vectors/s",
                    "ms_per_step": round(r1["ms_per_step"], 4),
                    "layout": r1["layout"],
                    "launch_shape": shape_json(r1["shape"])},
        "speedup_vs_fp64_k1": round(r32["value"] / r1["value"], 3),
    }
    tmax = min(r32["corr"].shape[1], r64["corr"].shape[1])
    out["ct_err_fp32_vs_fp64"] = float(np.max(np.abs(r32["corr"][:, :tmax] -
                                                     r64["corr"][:, :tmax])))
    ex = finish_exact(exact_proc) if exact_proc is not None else None
    if ex is not None:
        pts, cex = ex["dos"]
        pts = [p for p in pts if p < tmax]
        k = len(pts)
        out["ct_err_fp32_vs_exact"] = float(np.max(np.abs(r32["corr"][:, pts] - cex[:, :k])))
        out["ct_err_fp64_vs_exact"] = float(np.max(np.abs(r64["corr"][:, pts] - cex[:, :k])))
        out["ct_err_exact_steps"] = {"states": K, "run_steps": tmax - 1,
                                     "sampled_steps": len(pts), "last": pts[-1]}
        log(f"max |C - C_exact| over {tmax - 1} steps: fp32 "
            f"{out['ct_err_fp32_vs_exact']:.3e}, fp64 {out['ct_err_fp64_vs_exact']:.3e}")
    if cpu is not None:
        out["cpu_baseline"] = cpu
    return out, ex


# ---------------------------------------------------------------------------
# Other workloads (sub-records)
# ---------------------------------------------------------------------------
def cond_order(s, tile):
    return s.tile_order(tile, return_tiles=True) if tile > 0 else (None, None)


def ac_workload(args, exact=None):
    """Pristine graphene 2048x2048 (rescale 9), AC conductivity correlation
    of 8 states at 300 K, mu = 0, run as the Solver runs it: --ac-states
    states at a time (default 4) on the tiles of the library's AC rule for
    each precision (3.8 nm for complex64, 2.5 nm for complex128) unless
    --cond-tile.  Per run the Fermi and
    one-minus-Fermi polynomials (FFT coefficients at 2^15 points, precision
    1e-12) are applied once, then J_x psi1, J_y psi1 and psi2 of every state
    propagated together with tau = pi/2 (16 Bessel terms at 1e-14) and the
    four current correlations formed every step; complex64 (and complex128
    for the error) with FP64 sums.  state-steps/s = states x steps over the
    propagation time of all runs (the Fermi operators excluded)."""
    import numpy as np
    import torch
    import tbplas_amd as tb
    from tbplas_amd import tbpm, _lib, tune
    L, K = args.L, args.vectors
    per = max(1, min(args.ac_states, K))
    T = max(args.steps, 2)
    s = tb.make_graphene_sample(L)
    s.rescale_ham(9.0)
    ip, ix, hop, dx, dy = s.build_ham_dxy_device()
    n, nnz = s.num_orb, int(ix.numel())
    ts = 2 * math.pi / s.energy_range * 0.5
    bes = tbpm.bessel_series(ts * s.rescale, 1e-14)
    beta = 1.0 / (8.617333262e-5 * 300.0) * s.rescale
    cF = tbpm.fermi_cheb_coef(2 ** 15, beta, 0.0, False, 1e-12)
    cO = tbpm.fermi_cheb_coef(2 ** 15, beta, 0.0, True, 1e-12)
    h = _lib.default_handle(0)
    h.set_graph(not args.no_graph)
    tile_of = {prec: (round(tb.ac_tile_nm(3 * per, prec), 1) if args.cond_tile is None
                      else args.cond_tile) for prec in ("fp32", "fp64")}
    psi = tb.states.random_states(n, tb.states.sample_seeds(1337, K))
    chunks = [(a, min(K, a + per)) for a in range(0, K, per)]
    res = {}
    for prec in ("fp32", "fp64"):
        tile = tile_of[prec]
        order, tiles = cond_order(s, tile)
        H = tbpm.Hamiltonian(n, ip, ix, hop, dx, dy, cur_scale=s.rescale, precision=prec,
                             handle=h, order=order, tiles=tiles)
        real = H.info()["real_values"]
        shape = tune.apply(H, 3 * per, bes, not args.no_tune)
        threads = H.blocking_launch(3 * per)[1]
        sched = H.blocking_bytes(3 * per)
        t_prop, t_fermi, cs = 0.0, 0.0, []
        for a, b in chunks:
            # Fermi operators alone (T = 1 also forms one correlation step)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            tbpm.corr_ac(H, bes, cF, cO, psi[a:b], 1, chk_thr=1e-4)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            cs.append(tbpm.corr_ac(H, bes, cF, cO, psi[a:b], T, chk_thr=1e-4))
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            t_prop += (t2 - t1) - (t1 - t0)
            t_fermi += t1 - t0
        c = np.concatenate(cs)
        kt = {}
        if args.profile_steps > 0:
            # one run of the first chunk; the library leaves the Fermi
            # operators' launches untimed
            h.set_kernel_timing(True)
            try:
                a, b = chunks[0]
                tbpm.corr_ac(H, bes, cF, cO, psi[a:b], 1 + args.profile_steps, chk_thr=1e-4)
                kt = h.kernel_times()
            finally:
                h.set_kernel_timing(False)
        acc = None
        if exact is not None and L == args.L and K == exact[1].shape[0]:
            # 256-step runs against the exact k-space correlations
            pts, cex = exact
            ca = np.concatenate([tbpm.corr_ac(H, bes, cF, cO, psi[a:b], AC_ACC_STEPS,
                                              chk_thr=1e-4) for a, b in chunks])
            acc = float(np.max(np.abs(ca[:, :, [t - 1 for t in pts]] - cex)))
        H.close()
        sz = 8 if prec == "fp32" else 16

        def bytes_of(kind, sched=sched, sz=sz):
            if kind == "ac_corr":
                # cols, J_x, J_y values (complex), psi1x, psi1y gathered, psi2 own row
                return nnz * (4 + 2 * sz) + 3 * per * n * sz
            return pass_bytes_fn(sched, n, nnz, 3 * per, sz, real)(kind)
        label = (lambda k, prec=prec, real=real, threads=threads:
                 kernel_label(k, prec, 3 * per, real, threads))
        res[prec] = {"value": K * (T - 1) / t_prop, "ms_per_step": t_prop / (T - 1) * 1e3,
                     "fermi_setup_s": t_fermi, "shape": shape, "corr": c,
                     "layout": (f"tiles of {tile:g} nm" if tile > 0 else "caller order")
                     + passes_name(shape),
                     "roofline": (roofline(kt, bytes_of, label, traffic_key(f"ac_{prec}", tile))
                                  if kt else None),
                     "err_vs_exact": acc}
        log(f"ac {prec}: {res[prec]['value']:.1f} state·steps/s, max |C - C_exact| {acc}")
    err = float(np.max(np.abs(res["fp32"]["corr"] - res["fp64"]["corr"])))
    scale = float(np.max(np.abs(res["fp64"]["corr"])))
    r = res["fp32"]
    return {"metric": "AC correlation state·steps/s (8.4M-orbital graphene, 3 propagated "
                      "vectors per state)",
            "value": round(r["value"], 2), "unit": "state·steps/s", "higher_is_better": True,
            "dtype": "fp32", "ms_per_step": round(r["ms_per_step"], 4),
            "fermi_setup_s": round(r["fermi_setup_s"], 3),
            "config": {"workload": f"graphene {L}x{L} AC, {K} states in runs of {per} "
                                   f"({3 * per} vectors each), {T} steps, {len(bes)} Bessel "
                                   f"terms, Fermi terms {len(cF)}",
                       "n_orbitals": n, "nnz": nnz,
                       "layout": r["layout"],
                       "launch_shape": shape_json(r["shape"])},
            "roofline": r["roofline"],
            "fp64": {"value": round(res["fp64"]["value"], 2),
                     "ms_per_step": round(res["fp64"]["ms_per_step"], 4),
                     "layout": res["fp64"]["layout"],
                     "roofline": res["fp64"]["roofline"]},
            "corr_err_fp32_vs_fp64": err, "corr_max_abs": scale, "corr_err_steps": T,
            "ct_err_fp32_vs_exact": res["fp32"]["err_vs_exact"],
            "ct_err_fp64_vs_exact": res["fp64"]["err_vs_exact"],
            "ct_err_exact_steps": {"run_steps": AC_ACC_STEPS, "sampled": AC_PTS,
                                   "states": K} if exact is not None else None}


def dc_workload(args):
    """Graphene 1024x1024, 1% vacancies, Anderson disorder U[-0.5, 0.5] eV,
    Peierls phases for B = 20 T (gauge 0), rescale 9; DC correlation at the 5
    grid energies nearest -1, -0.5, 0, 0.5, 1 eV; 16 states, --steps steps
    (at most 1024) in runs of --dc-states, complex64 (FP64 quasi-eigenstates
    and correlations), tiles sized by the library's rule for a run's
    vectors (3.2 nm for 8 states) unless --cond-tile."""
    import numpy as np
    import torch
    import tbplas_amd as tb
    from tbplas_amd import tbpm, _lib, tune
    L, K = 1024, 16
    per = max(1, min(args.dc_states, K))
    T = max(2, min(args.steps, 1024))
    tile = (round(tb.auto_tile_nm(2 * per, "fp32"), 1) if args.cond_tile is None
            else args.cond_tile)
    s = tb.make_graphene_sample(L, vacancies=tb.random_vacancies(2 * L * L, 0.01, 2024))
    tb.add_anderson_disorder(s, 1.0, 2025)
    s.set_magnetic_field(20.0, 0)
    s.rescale_ham(9.0)
    ip, ix, hop, dx, dy = s.build_ham_dxy_device()
    n, nnz = s.num_orb, int(ix.numel())
    en = np.array([0.5 * i * s.energy_range / T - s.energy_range / 2. for i in range(2 * T)])
    inds = np.array([int(np.argmin(np.abs(en - e))) for e in (-1.0, -0.5, 0.0, 0.5, 1.0)])
    ts = 2 * math.pi / s.energy_range
    bes = tbpm.bessel_series(ts * s.rescale, 1e-14)
    h = _lib.default_handle(0)
    h.set_graph(not args.no_graph)
    order, tiles = cond_order(s, tile)
    H = tbpm.Hamiltonian(n, ip, ix, hop, dx, dy, cur_scale=s.rescale, precision="fp32",
                         handle=h, order=order, tiles=tiles)
    real = H.info()["real_values"]
    shape = tune.apply(H, 2 * per, bes, not args.no_tune)
    threads = H.blocking_launch(2 * per)[1]
    sched = H.blocking_bytes(2 * per)
    psi = tb.states.random_states(n, tb.states.sample_seeds(1337, K))
    chunks = [(a, min(K, a + per)) for a in range(0, K, per)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    parts = [tbpm.corr_dc(H, bes, psi[a:b], T, ts, en, inds, chk_thr=1e-4) for a, b in chunks]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    d = np.concatenate([p[0] for p in parts])
    c = np.concatenate([p[1] for p in parts])
    kt = {}
    if args.profile_steps > 0:
        h.set_kernel_timing(True)
        try:
            Tp = max(2, min(args.profile_steps, T))
            enp = np.array([0.5 * i * s.energy_range / Tp - s.energy_range / 2.
                            for i in range(2 * Tp)])
            indp = np.array([int(np.argmin(np.abs(enp - e))) for e in (-1.0, -0.5, 0.0, 0.5, 1.0)])
            tbpm.corr_dc(H, bes, psi[chunks[0][0]:chunks[0][1]], Tp, ts, enp, indp,
                         chk_thr=1e-4)
            kt = h.kernel_times()
        finally:
            h.set_kernel_timing(False)
    H.close()
    ne = len(inds)

    def bytes_of(kind):
        if kind == "dc_qe":    # psi(+t), psi(-t), psi0 read; QE read and written (FP64)
            return 3 * per * n * 8 + 2 * ne * per * n * 16
        if kind == "dc_dot":   # J_a psi(t) read, J_a QE read (FP64)
            return 2 * per * n * 8 + 2 * ne * per * n * 16
        return pass_bytes_fn(sched, n, nnz, 2 * per, 8, real)(kind)
    value = 2 * K * (2 * T - 1) / dt
    return {"metric": "DC correlation step·vectors/s (disordered field-threaded graphene)",
            "value": round(value, 2), "unit": "step·vectors/s", "higher_is_better": True,
            "dtype": "fp32", "seconds": round(dt, 3),
            "config": {"workload": f"graphene {L}x{L}, 1% vacancies, Anderson W=1 eV, B=20 T, "
                                   f"DC at 5 energies, {K} states in runs of {per}, {T} steps",
                       "n_orbitals": n, "nnz": nnz, "energies": en[inds].tolist(),
                       "layout": (f"tiles of {tile:g} nm" if tile > 0 else "caller order")
                       + passes_name(shape),
                       "launch_shape": shape_json(shape)},
            "roofline": (roofline(kt, bytes_of,
                                  lambda k: kernel_label(k, "fp32", 2 * per, real, threads),
                                  traffic_key("dc", tile))
                         if kt else None),
            "dos_corr_t1_mean": float(np.mean(d[:, 0]).real),
            "dc_corr_t1_mean": [float(np.mean(c[:, a, :, 0]).real) for a in range(2)]}


def batch_workload(args):
    """32 independent disordered graphene instances of uneven size (edges from
    {128, 256, 384, 512, 768, 1024}, vacancies 0-3%, Anderson W 0-2 eV, all
    seeded), rescaled per instance by the row-sum bound, 4 states each, DOS
    correlation for --steps steps (at most 1024), every instance in one
    launch per pass (rows in --tile nm tiles), replayed from a hipGraph;
    complex64, with the per-instance error against the complex128 run on
    the same states."""
    import numpy as np
    import tbplas_amd as tb
    from tbplas_amd import tbpm, _lib, tune
    rng = np.random.default_rng(4)
    edges = rng.choice([128, 256, 384, 512, 768, 1024], 32)
    vac = rng.uniform(0.0, 0.03, 32)
    Ws = rng.uniform(0.0, 2.0, 32)
    K = 4
    tiles_on = args.order == "tiles"
    tile = leg_tile(args, K, "fp32") if tiles_on else 0
    tile64 = leg_tile(args, K, "fp64") if tiles_on else 0
    h = _lib.default_handle(0)
    h.set_graph(not args.no_graph)
    insts = []
    for i in range(32):
        L = int(edges[i])
        s = tb.make_graphene_sample(L, vacancies=tb.random_vacancies(2 * L * L, vac[i], 100 + i))
        tb.add_anderson_disorder(s, Ws[i], 200 + i)
        s.rescale_ham()
        ip, ix, dt = s.build_ham_csr_device()
        insts.append((s.num_orb, int(ix.numel()), ip, ix, dt,
                      tb.states.random_states(s.num_orb, tb.states.sample_seeds(300 + i, K)),
                      {"fp32": s.tile_order(tile, return_tiles=True) if tile else (None, None),
                       "fp64": (s.tile_order(tile64, return_tiles=True) if tile64
                                else (None, None))}))
    bes = tbpm.bessel_series(math.pi, 1e-14)
    total = args.warmup + args.steps + args.profile_steps
    res = {}
    for prec in ("fp32", "fp64"):
        hams = [tbpm.Hamiltonian(n, ip, ix, dt, precision=prec, handle=h, order=ot[prec][0],
                                 tiles=ot[prec][1])
                for (n, nnz, ip, ix, dt, _, ot) in insts]
        shape = tune.apply(hams, K, bes, not args.no_tune)
        real = all(H.info()["real_values"] for H in hams)
        run = tbpm.DosRun(hams, bes, [x[5] for x in insts], total)
        ms, kt = timed_run(run, h, args, total)
        res[prec] = (ms, run.result(chk_step=0), shape, real, kt, run.blocking_bytes())
        run.close()
        for H in hams:
            H.close()
    ms, c32, shape, real, kt, sched = res["fp32"]
    c64 = res["fp64"][1]
    errs = [float(np.max(np.abs(a - b))) for a, b in zip(c32, c64)]
    n_tot = sum(x[0] for x in insts)
    nnz_tot = sum(x[1] for x in insts)
    t_step = ms / 1e3 / args.steps
    return {"metric": "uneven batch DOS step·vectors/s (32 instances in one launch per pass)",
            "value": round(32 * K / t_step, 2), "unit": "instance-step·vectors/s",
            "higher_is_better": True, "dtype": "fp32", "ms_per_step": round(t_step * 1e3, 4),
            "config": {"workload": f"32 disordered graphene instances, edges "
                                   f"{sorted(set(edges.tolist()))}, {K} states each, "
                                   f"{args.steps} timed steps", "n_orbitals_total": n_tot,
                       "nnz_total": nnz_tot, "launch_shape": shape_json(shape),
                       "layout": (f"tiles of {tile:g} nm" if tile
                                  else "caller order") + passes_name(shape),
                       "layout_fp64": (f"tiles of {tile64:g} nm" if tile64 else "caller order")
                       + passes_name(res["fp64"][2]),
                       "hipgraph": not args.no_graph},
            "roofline": (roofline(kt, pass_bytes_fn(sched, n_tot, nnz_tot, K, 8, real),
                                  lambda k: kernel_label(k, "fp32", K, real, 1024),
                                  traffic_key("batch", tile))
                         if kt else None),
            "fp64_ms_per_step": round(res["fp64"][0] / args.steps, 4),
            "ct_err_fp32_vs_fp64_per_instance_max": max(errs),
            "ct_err_fp32_vs_fp64_per_instance_median": float(np.median(errs))}


def latency_workload(args):
    """Pristine graphene 256x256 (131,072 orbitals, rescale 9), one state,
    1024 DOS steps (21 Bessel terms at 1e-15): the case whose vectors fit in
    the last-level cache, reported as microseconds per step, complex128 and
    complex64, hipGraph replay, tiles by the library's rule for a small
    sample (two per compute unit: 2.6 nm on MI355X; the tuner picks one,
    three, four or five terms per pass)."""
    import torch
    import tbplas_amd as tb
    from tbplas_amd import tbpm, _lib, tune
    L, T = 256, 1024
    s = tb.make_graphene_sample(L)
    s.rescale_ham(9.0)
    ip, ix, dt = s.build_ham_csr_device()
    n, nnz = s.num_orb, int(ix.numel())
    bes = tbpm.bessel_series(2 * math.pi / s.energy_range * s.rescale, 1e-15)
    psi = tb.states.random_states(n, tb.states.sample_seeds(1337, 1))
    h = _lib.default_handle(0)
    h.set_graph(not args.no_graph)
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    tile = round(tb.auto_tile_nm(1, "fp64", n_orb=n, cus=cus), 1)
    order, tiles = (s.tile_order(tile, return_tiles=True) if args.order == "tiles"
                    else (None, None))
    a = argparse.Namespace(**vars(args))
    a.warmup, a.steps = 32, T
    res = {}
    for prec in ("fp64", "fp32"):
        H = tbpm.Hamiltonian(n, ip, ix, dt, precision=prec, handle=h, order=order, tiles=tiles)
        real = H.info()["real_values"]
        shape = tune.apply(H, 1, bes, not args.no_tune)
        threads = H.blocking_launch(1)[1]
        run = tbpm.DosRun(H, bes, psi, a.warmup + T + a.profile_steps)
        ms, kt = timed_run(run, h, a, a.warmup + T + a.profile_steps)
        sched = run.blocking_bytes()
        run.close()
        H.close()
        sz = 8 if prec == "fp32" else 16
        rl = (roofline(kt, pass_bytes_fn(sched, n, nnz, 1, sz, real),
                       lambda k, prec=prec, real=real, threads=threads:
                       kernel_label(k, prec, 1, real, threads),
                       traffic_key(f"latency_{prec}", tile if order is not None else 0))
              if kt else None)
        if rl is not None:
            rl["note"] = ("the vectors fit in the 256 MB last-level cache: launch and "
                          "latency bound, HBM bandwidth does not bound it")
        res[prec] = {"us_per_step": round(ms / T * 1e3, 2),
                     "us_per_term": round(ms / T / (len(bes) - 1) * 1e3, 3),
                     "layout": ("caller order" if order is None else f"tiles of {tile:g} nm")
                     + passes_name(shape),
                     "launch_shape": shape_json(shape), "roofline": rl}
    return {"metric": "TBPM microseconds per step, 131,072-orbital graphene, 1 state",
            "value": res["fp64"]["us_per_step"], "unit": "us/step", "higher_is_better": False,
            "dtype": "fp64",
            "config": {"workload": f"graphene {L}x{L} DOS, 1 state, {T} steps",
                       "n_orbitals": n, "nnz": nnz, "hipgraph": not args.no_graph},
            "fp64": res["fp64"], "fp32": res["fp32"]}


def assembly_workload(args):
    """GPU Hamiltonian assembly from the hop list: the insertion-order CSR
    with dx, dy (build_ham_dxy_fast's output, sample.pyx:148-257) and the
    canonical CSR (build_ham_csr) at 8.4M orbitals (2048x2048 graphene) and
    10^7 atoms (2236x2236), timed with the inputs already in device memory
    (count + fill) and end to end from host arrays to host arrays."""
    import ctypes
    import torch
    import tbplas_amd as tb
    from tbplas_amd import _lib
    from tbplas_amd.sample import _DevHops
    h = _lib.default_handle(0)
    L_ = _lib.lib()
    recs = {}
    for L in (2048, 2236):
        s = tb.make_graphene_sample(L)
        s.rescale_ham(9.0)
        s.init_orb_pos()
        n, m = s.num_orb, s.num_hop
        d = _DevHops(s, 0, with_dr=True)
        dev = "cuda:0"

        def dxy():
            indptr = torch.empty(n + 1, dtype=torch.int64, device=dev)
            nnz = ctypes.c_int64(0)
            _lib.check(L_.tbg_ham_dxy_count(h.ptr, n, m, _lib.ptr(d.orb_eng), _lib.ptr(d.hop_i),
                                            _lib.ptr(d.hop_j), 1e-5, _lib.ptr(indptr),
                                            ctypes.byref(nnz)))
            k = nnz.value
            ix = torch.empty(k, dtype=torch.int64, device=dev)
            hv = torch.empty(k, dtype=torch.complex128, device=dev)
            dx = torch.empty(k, dtype=torch.float64, device=dev)
            dy = torch.empty(k, dtype=torch.float64, device=dev)
            _lib.check(L_.tbg_ham_dxy_fill(h.ptr, n, m, _lib.ptr(d.orb_eng), _lib.ptr(d.hop_i),
                                           _lib.ptr(d.hop_j), _lib.ptr(d.hop_v), _lib.ptr(d.dr),
                                           1e-5, _lib.ptr(indptr), _lib.ptr(ix), _lib.ptr(hv),
                                           _lib.ptr(dx), _lib.ptr(dy)))
            return k

        def csr():
            indptr = torch.empty(n + 1, dtype=torch.int64, device=dev)
            nnz = ctypes.c_int64(0)
            _lib.check(L_.tbg_ham_csr_count(h.ptr, n, m, _lib.ptr(d.orb_eng), _lib.ptr(d.hop_i),
                                            _lib.ptr(d.hop_j), _lib.ptr(d.hop_v),
                                            _lib.ptr(indptr), ctypes.byref(nnz)))
            k = nnz.value
            ix = torch.empty(k, dtype=torch.int64, device=dev)
            dt = torch.empty(k, dtype=torch.complex128, device=dev)
            _lib.check(L_.tbg_ham_csr_fill(h.ptr, n, m, _lib.ptr(d.orb_eng), _lib.ptr(d.hop_i),
                                           _lib.ptr(d.hop_j), _lib.ptr(d.hop_v),
                                           _lib.ptr(indptr), _lib.ptr(ix), _lib.ptr(dt)))
            return k

        def best(f, reps=5):
            f()
            torch.cuda.synchronize()
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                k = f()
                torch.cuda.synchronize()
                ts.append(time.perf_counter() - t0)
            return min(ts), k
        t_dxy, nnz_dxy = best(dxy)
        t_csr, nnz_csr = best(csr)
        t0 = time.perf_counter()
        s.build_ham_dxy()
        t_host = time.perf_counter() - t0
        recs[f"{n}_orbitals"] = {
            "L": L, "n_orbitals": n, "hops": m, "nnz": nnz_dxy,
            "dxy_device_s": round(t_dxy, 5), "csr_device_s": round(t_csr, 5),
            "dxy_host_to_host_s": round(t_host, 4),
            "dxy_entries_per_s": round(nnz_dxy / t_dxy, 1), "csr_nnz": nnz_csr}
        del d
        torch.cuda.empty_cache()
    return {"metric": "GPU Hamiltonian assembly time (insertion-order CSR + dx, dy; canonical "
                      "CSR)", "unit": "s", "higher_is_better": False, "dtype": "int64+fp64",
            "value": recs["9999392_orbitals"]["dxy_device_s"],
            "reference_published": {"build_ham_dxy_fast_s": 1.48, "build_ham_dxy_safe_s": 2.83,
                                    "atoms": "1e7", "source": "tbplas/cython/sample.pyx:299-301"},
            "sizes": recs}


def hall_workload(args):
    """The reference's Hall example (tests/test_units/test_hall.py): graphene
    500x200 with t = 3 eV (200,000 orbitals), B = 100 T (gauge 0), rescale 9,
    Kubo-Bastin moments with 2048 Chebyshev kernels in direction x for one
    state, complex128.  The n_kernel^2 Gram matrix runs on the FP64 matrix
    cores (k_gram, eight blocks of 256 rows): its roofline is FP64 MFMA,
    8 flops per complex multiply-add over the sample's orbitals."""
    import torch
    import tbplas_amd as tb
    from tbplas_amd import tbpm, _lib
    N = 2048
    s = tb.make_graphene_sample(500, 200, t=3.0)
    s.set_magnetic_field(100.0, 0)
    s.rescale_ham(9.0)
    ip, ix, hop, dx, dy = s.build_ham_dxy_device()
    n = s.num_orb
    h = _lib.default_handle(0)
    H = tbpm.Hamiltonian(n, ip, ix, hop, dx, dy, cur_scale=s.rescale, precision="fp64",
                         handle=h)
    psi = tb.states.random_states(n, tb.states.sample_seeds(1337, 1))
    tbpm.hall_mu(H, psi, 256, 1)          # code objects and buffers warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mu = tbpm.hall_mu(H, psi, N, 1)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    kt = {}
    if args.profile_steps > 0:
        h.set_kernel_timing(True)
        try:
            mu2 = tbpm.hall_mu(H, psi, N, 1)
            kt = h.kernel_times()
        finally:
            h.set_kernel_timing(False)
        assert (mu2 == mu).all(), "Hall moments differ between two runs"
    H.close()
    rl = None
    if "gram" in kt:
        ms, cnt = kt["gram"]
        avg_s = ms / cnt / 1e3
        flops = 8.0 * 256 * N * n          # one block of 256 rows of mu
        ach = flops / avg_s / 1e12
        # only the Gram blocks are timed per launch here: their share of
        # the untimed run's wall time
        rl = {"bound": "mfma", "kernel": "k_gram", "launches": cnt,
              "launch_avg_us": round(avg_s * 1e6, 2),
              "share_of_run_time": round(ms / 1e3 / dt, 4),
              "achieved": round(ach, 2), "peak": FP64_PEAK_TFS, "unit": "TFLOP/s",
              "frac": round(ach / FP64_PEAK_TFS, 4), "flops_per_launch": flops,
              "traffic": None,
              "launches_by_kind": {k: {"launches": v[1], "avg_us": round(v[0] / v[1] * 1e3, 2)}
                                   for k, v in sorted(kt.items(), key=lambda kv: -kv[1][0])}}
        tr = counter_traffic("hall", "k_gram")
        if tr is not None:
            rl["traffic"] = tr["bytes_per_launch"]
            rl["traffic_source"] = tr.get("summary")
    return {"metric": "KPM Hall moments, seconds for 2048 kernels (200,000-orbital graphene, "
                      "B = 100 T)", "value": round(dt, 4), "unit": "s", "higher_is_better": False,
            "dtype": "fp64",
            "config": {"workload": f"graphene 500x200, t = 3 eV, B = 100 T, n_kernel {N}, "
                                   "direction x, 1 state", "n_orbitals": n,
                       "nnz": int(ix.numel())},
            "roofline": rl, "mu_00": [float(mu[0, 0, 0].real), float(mu[0, 0, 0].imag)]}


def main():
    args = parse()
    if args.cpu_child:
        cpu_child(args)
        return 0
    if args.exact_child:
        exact_child(args)
        return 0
    if args.gpus != 1:
        print(f"bench.py drives one GPU; --gpus {args.gpus} is not supported "
              "(states are independent: several GPUs would each propagate a share "
              "of them and sum the correlation partials once per run).", file=sys.stderr)
        return 2
    if args.copy_probe:
        copy_probe()
        return 0
    subs = {"ac": ac_workload, "dc": dc_workload, "batch": batch_workload,
            "latency": latency_workload, "assembly": assembly_workload, "hall": hall_workload}
    if args.workload in subs:
        rec = subs[args.workload](args)
        rec["n_gpus"] = 1
        print(json.dumps(rec), flush=True)
        return 0
    cpu = None if args.no_cpu else cpu_baseline(args)
    exact = start_exact(args) if args.accuracy_steps >= ACC_STEPS else None
    out, ex = headline(args, cpu, exact)
    if args.workload == "all":
        for name, fn in subs.items():
            t0 = time.perf_counter()
            out[name] = fn(args, ex["ac"] if ex is not None else None) if name == "ac" else fn(args)
            log(f"{name}: {out[name]['value']} {out[name]['unit']} "
                f"({time.perf_counter() - t0:.1f} s)")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
