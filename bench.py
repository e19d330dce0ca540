"""NNP molecular-liquid MD benchmark on one MI355X.

Workload (24k-atom water NNP step): 8000 flexible 3-site water molecules (24000 atoms) in a
62.1 A box; tabulated LJ + shifted-force Coulomb to 9 A (0.01 A table); symmetry-function NNP to
6 A (16 radial + 16 angular per element pair, D = 80) with a per-element 3 x 128 tanh MLP;
harmonic bonds/angles; full lists to 11 A (2 A skin) rebuilt every 10 steps; O-O/O-H/H-H g(r)
on 0-12 A in 0.02 A bins and S(q) on 1000 q in 0.3-20 1/A every 10 steps; FP64 velocity Verlet
at dt = 0.5 fs.  A step is one MD step of this workload; steps replay captured hipGraphs.

value = atom-steps/s of the BF16-MFMA form (whole job, one GPU, inputs resident in HBM); the
FP32 form runs the same steps with FP32 network arithmetic.  Force errors are relative RMS
against the FP64 oracle on the starting snapshot.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NMOL, BL = 8000, 62.1
RC_PAIR, SKIN, RC_NNP, R_GR = 9.0, 2.0, 6.0, 12.0
WIDTHS = [80, 128, 128, 128, 1]
DT = 0.5
HBM_PEAK_GBS = 8000.0
MFMA_PEAK = {"bf16": 2500.0, "fp16": 2500.0, "fp32": 157.3}  # TFLOP/s dense
# non-packed FP32 vector rate: 256 CUs x 64 lanes x 2 (FMA) x 2.4 GHz (the library builds without
# packed FP32 instructions)
VALU_PEAK_TFLOPS = 78.6
CLOCK_HZ = 2.4e9  # MI355X peak engine clock
COPY_PROBE_BYTES = 64 << 20  # the copy stage: 64 MiB read and 64 MiB written
MFMA_PROBE_FLOP = 2048 * 2048 * 4 * 2 * 32 * 32 * 16  # the mfma stage: waves x rounds x chains x flop per MFMA
# algorithmic FP32 operations per directed pair within the symmetry-function cutoff
SF_FORCE_FLOP_PAIR = 405   # geometry 16, cutoff 32, angular weights 21, radial polynomials 74,
                           # adjoint combination 60, angular gradient 4 x 48, force 10
SF_DESC_FLOP_PAIR = 205    # geometry 16, cutoff 18, radial Gaussians and sums 67, angular moments 4 x 26


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_engine(sysm, par, norm, prec):
    from nnpengine_amd import model as M
    from nnpengine_amd import system as S
    from nnpengine_amd.engine import Engine
    ft, ut = S.lj_table(sysm, RC_PAIR, 900, S.WATER_LJ)
    sfp = M.sf_params(RC_NNP)
    e = Engine(sysm)
    e.neigh_init(RC_PAIR, SKIN, R_GR, RC_NNP)
    e.pair_table(RC_PAIR, ft, ut)
    e.bonded_set_from_system()
    e.sf_set(sfp["rc"], sfp["mu"], sfp["eta_r"], sfp["eta_a"])
    e.mlp_set(WIDTHS, par, 0.5, prec, *norm)
    e.gr_init(450, R_GR, RC_PAIR)
    e.sq_init(1000, 0.3, 20.0)
    return e


def gpu_standardisation(sysm):
    """Per-element input standardisation of the network from the starting snapshot's
    descriptors, computed by the engine."""
    from nnpengine_amd import model as M
    from nnpengine_amd.engine import Engine
    sfp = M.sf_params(RC_NNP)
    e = Engine(sysm)
    e.neigh_init(RC_PAIR, SKIN, R_GR, RC_NNP)
    e.sf_set(sfp["rc"], sfp["mu"], sfp["eta_r"], sfp["eta_a"])
    e.build()
    e.sf_descriptor()
    G0 = e.nnp_get("G").astype(np.float64)
    e.close()
    return M.standardisation(G0, sysm.typ, sysm.ntype)


MD_KW = dict(dt=DT, fp64=True, rebuild_every=10, pair=True, bonded=True, nnp=2, gr_every=10, gr_long_every=10,
             sq=True, use_graph=True)


def time_steps(e, steps, warmup):
    e.md_run(warmup, log_every=0, **MD_KW)
    e.sync()
    t0 = time.perf_counter()
    e.md_run(steps, log_every=0, **MD_KW)
    e.sync()
    return time.perf_counter() - t0


def cpu_reference(sysm):
    """The reference's own classical MD loop, compiled from its sources (oracle/_ref/ref_harness:
    refold/sort/ghosts/NEIGHLIST::build on NEIGHLIST::isvalid, BOND/ANGLE/DIHEDRAL::compute,
    LJTableCoulDSF::compute, VERLET, src/run.cpp:204-394) with its 27 OpenMP threads
    (src/environ.h:85-90), on the same 24000-atom snapshot and velocities at dt = 0.5 fs.  The
    time per step is the difference of the MD-loop wall times of a 10-step and a 30-step run.
    Returns None when the harness is not built."""
    from nnpengine_amd import system as S
    from oracle import refrun
    if not refrun.available():
        return None
    ft, ut = S.lj_table(sysm, RC_PAIR, 900, S.WATER_LJ)
    t = {}
    for n in (10, 30):
        r = refrun.run(sysm, RC_PAIR, SKIN, RC_PAIR, ft, ut, mode="md", nsteps=n, dt=DT)
        t[n] = (r["t_loop"], int(r["nbuild"]), int(r["threads"]))
    per_step = (t[30][0] - t[10][0]) / 20
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count()
    threads = t[30][2]
    return {"value": sysm.natom / per_step, "unit": "atom-steps/s", "cores": cores, "threads": threads, "kind": "reference",
            "ms_per_step": per_step * 1e3,
            "oversubscribed": threads > cores,
            "sample": f"the reference's classical step compiled from its sources, {threads} OpenMP threads on {cores} "
                      f"host cores: tabulated LJ + shifted-force Coulomb 9 A, bonded terms, lists 11 A rebuilt when an "
                      f"atom moved > skin/2 ({t[30][1]} builds in 30 steps), FP32 Verlet, 24000-atom water at dt 0.5 fs; "
                      f"per-step time from 10- and 30-step runs ({t[10][0]:.2f} s, {t[30][0]:.2f} s).  The NNP and g(r) "
                      f"are not in it: pair_nnp.cpp and grsq.cpp include the TensorFlow C header, absent here"}


def cpu_port(sysm, par, norm):
    """The oracle's float32 restatement of the whole NNP step on the host (one thread), on a
    bounded sample: 5 steps of the 24k-atom workload (about 10 s)."""
    from nnpengine_amd import model as M
    from nnpengine_amd import system as S
    from oracle import pyoracle
    ft, ut = S.lj_table(sysm, RC_PAIR, 900, S.WATER_LJ)
    sfp = M.sf_params(RC_NNP)
    o = pyoracle.Oracle(sysm, 0)
    o.neigh_init(RC_PAIR, SKIN, R_GR)
    o.pair_set(RC_PAIR, ft, ut)
    o.sf_set(sfp["rc"], sfp["mu"], sfp["eta_r"], sfp["eta_a"])
    o.mlp_set(WIDTHS, par, 0.5, *norm)
    o.gr_init(450, R_GR, RC_PAIR)
    o.set_dt(DT)
    nsteps = 5
    t_reb = t_gr = t_step = 0.0
    for s in range(nsteps):
        t0 = time.perf_counter()
        if s > 0:
            o.update_x()
        o.rebuild(True)
        t1 = time.perf_counter()
        o.gr_tally(True)
        o.sq(1000, 0.3, 20.0)
        t2 = time.perf_counter()
        o.clear_f()
        o.bonded_compute(False)
        o.pair_compute(False)
        G = o.sf_descriptor()
        _, _, dG = o.mlp_eval(G)
        o.sf_force(dG, False)
        o.update_v()
        t3 = time.perf_counter()
        t_reb += t1 - t0
        t_gr += t2 - t1
        t_step += t3 - t2
    # the workload rebuilds lists and tallies g(r)/S(q) every 10th step
    per_step = (t_step + (t_reb + t_gr) / 10) / nsteps
    total = t_reb + t_gr + t_step
    return {"value": sysm.natom / per_step, "unit": "atom-steps/s", "cores": 1, "threads": 1, "kind": "port",
            "sample": f"{nsteps} steps of the 24000-atom workload ({total:.1f} s): forces (pair, bonded, NNP "
                      f"descriptors, FP64 MLP, NNP forces) + Verlet every step, list rebuild and 12 A g(r) + S(q) "
                      f"charged at 1/10 per step; oracle float32 restatement, 1 thread"}


def _serve_box(args):
    from nnpengine_amd import system as S
    nmol, seed = args
    return S.water_box(nmol, 62.1 * (nmol / 8000.0) ** (1.0 / 3.0), seed=seed)


def serving(par, norm, nreq, steps, seed=2027):
    """Serving form: nreq independent water boxes (molecule counts log-uniform in [64, 4096],
    seeded, density of the 24k box), each `steps` NVE steps of the same BF16 NNP + LJ/Coulomb +
    bonded step, arriving as a Poisson stream at 80 % of the measured serving capacity; every
    request's latency runs from its arrival to the return of its final state."""
    from concurrent.futures import ProcessPoolExecutor
    from nnpengine_amd import system as S
    from nnpengine_amd.engine import Serve
    rng = np.random.default_rng(seed)
    nmols = np.exp(rng.uniform(np.log(64), np.log(4096), nreq)).astype(int)
    with ProcessPoolExecutor(max_workers=8) as ex:
        boxes = list(ex.map(_serve_box, [(int(n), seed + k) for k, n in enumerate(nmols)]))
    ref = boxes[0]
    srv = Serve(ref, WIDTHS, par, norm[0], norm[1], precision=HEAD, max_atoms=400000, max_boxes=256,
                pair_cutoff=RC_PAIR, skin=SKIN, nnp_cutoff=RC_NNP, rebuild_every=10, dt=DT, fp64=True)
    natom = np.array([b.natom for b in boxes])
    work = natom * steps
    # capacity: a closed batch of the first 32 requests, all submitted at once
    t0 = time.perf_counter()
    ids = [srv.submit(b, steps) for b in boxes[:32]]
    left = set(ids)
    while left:
        for rid in srv.advance():
            srv.result(rid)
            left.discard(rid)
    cap = work[:32].sum() / (time.perf_counter() - t0)
    # open loop: Poisson arrivals at 80 % of capacity
    rate = 0.8 * cap / work.mean()
    arrive = np.cumsum(rng.exponential(1.0 / rate, nreq))
    sub_of = {}
    lat = np.zeros(nreq)
    done = 0
    k = 0
    peak = 0
    t0 = time.perf_counter()
    while done < nreq:
        now = time.perf_counter() - t0
        while k < nreq and arrive[k] <= now:
            sub_of[srv.submit(boxes[k], steps)] = k
            k += 1
        st = srv.stats()
        peak = max(peak, st["resident"])
        if st["resident"] == 0 and st["queued"] == 0:
            time.sleep(max(0.0, min(arrive[k] - now, 0.01)) if k < nreq else 0.0)
            continue
        for rid in srv.advance():
            srv.result(rid)
            j = sub_of.pop(rid)
            lat[j] = time.perf_counter() - t0 - arrive[j]
            done += 1
    span = time.perf_counter() - t0 - arrive[0]
    st = srv.stats()
    srv.close()
    return {"requests": nreq, "steps_per_request": steps, "atoms_min": int(natom.min()), "atoms_max": int(natom.max()),
            "atoms_mean": float(natom.mean()), "capacity_atom_steps_per_s": float(cap), "arrival_rate_per_s": float(rate),
            "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                          "max": float(lat.max())},
            "atom_steps_per_s": float(work.sum() / span), "peak_resident_boxes": int(peak),
            "relayouts": int(st["relayouts"]), "chunks": int(st["chunks"]), "arena_bytes": int(st["arena_bytes"])}


E_TERMS = (0, 1, 2, 3, 4, 5, 7)  # pair, bond, angle, dihedral, NNP table pair, network, kinetic


def nve_drift(sysm, par, norm, prec, ps, log_every=200):
    """NVE energy drift of one precision form: the workload's MD step (FP64 Verlet) for `ps`
    picoseconds from the starting snapshot, total energy logged every `log_every` steps; the
    drift is the slope of a linear fit of E_tot / N against time.  Returns the drift record and
    the final positions."""
    e = make_engine(sysm, par, norm, prec)
    steps = int(round(ps * 1000 / DT))
    steps -= steps % log_every
    t0 = time.perf_counter()
    # chunks of 1000 steps; a chunk that overflows a list as the lattice melts is repeated from
    # its start with the grown capacities
    log, redo = e.md_run_checkpointed(steps, chunk=1000, log_every=log_every, **MD_KW)
    wall = time.perf_counter() - t0
    st = e.md_stats()
    x = e.get("x")
    e.close()
    et = log[:, [2 * k for k in E_TERMS]].sum(1) / sysm.natom
    t_ns = np.arange(len(et)) * log_every * DT * 1e-6
    # the step-0 record holds the starting velocities, which the integrator's start-up takes as
    # half-step velocities (the reference's setDt(-1) / updateX start, src/run.cpp:193): its
    # kinetic energy is not on the trajectory's energy surface, so the fit starts at record 1
    slope = float(np.polyfit(t_ns[1:], et[1:], 1)[0]) if len(et) > 3 else float("nan")
    ke = log[:, 14] / sysm.natom
    return {"ps": steps * DT / 1000, "steps": steps, "drift_kcal_mol_atom_per_ns": slope,
            "e_tot_per_atom_first_last": [float(et[1]), float(et[-1])], "e_tot_per_atom_std": float(et[1:].std()),
            "e_tot_per_atom_step0": float(et[0]),
            "ke_per_atom_mean": float(ke.mean()), "skin_violations": int(st["skin_violations"]),
            "repeated_chunks": int(redo), "wall_s": wall}, x


def sqd_timing(e, q_max=20.0, nq=1000, reps=3):
    """End-of-run analysis: direct k-space S(q) of the current 24000-atom positions over every
    reciprocal vector of the box with |G| <= 20 1/A (half space), 1000 bins."""
    t0 = time.perf_counter()
    ngv = e.sqd_init(nq, q_max)
    t_init = time.perf_counter() - t0
    e.sqd_compute()
    e.sync()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        e.sqd_compute()
        e.sync()
        ts.append(time.perf_counter() - t0)
    t = min(ts)
    _, sq, cnt = e.sqd_get()
    log(f"[bench] sqd: {ngv} vectors, {t * 1e3:.1f} ms per compute, init {t_init:.2f} s")
    return {"q_max": q_max, "nq": nq, "n_gvec": int(ngv), "ms": t * 1e3, "init_s": t_init,
            "atom_gvec_per_s": e.natom * ngv / t, "finite": bool(np.all(np.isfinite(sq[cnt > 0])))}


def hbm_traffic(stage):
    """HBM bytes per launch of the stage's kernel from the committed FETCH_SIZE/WRITE_SIZE summary
    (two rocprofv3 --pmc runs of this workload, scaled by a 64 MiB copy kernel's readings; the
    file states both commands and the scale factors), or None when it lists no such stage."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "pmc_hbm_traffic_water24k_nnp_step.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 6 and cells[1] == stage:
            try:
                return float(cells[5])
            except ValueError:
                return None
    return None


# the kernel each one-kernel stage launches in this workload (the rows of the VALU summary)
# the timed BF16 form: one BF16 term per weight, activation and gradient ("bf16x1"); the split
# hi + lo form ("bf16") is timed beside it
HEAD = "bf16x1"

STAGE_KERNEL = {"pair": "k_pair<false, 32, true>", "bonded": "k_bonded<false>", "sf_desc": "k_sf_desc<2, 16, true>",
                "mlp": "k_mlp_w<3, 3>", "sf_adjoint": "k_sf_adjoint<2, true>",
                "sf_force": "k_sf_force<2, false, 16, true, false, true>",
                "verlet": "k_update_x64"}


def active_pairs(e, bl, rc):
    """Entries of the short lists within the symmetry-function cutoff (the active lists the SF
    force pass reads): the FP32 distance test on the exported entries and positions."""
    ent = e.list_entries(1)
    x = e.get("x").astype(np.float32)
    i, j, code = ent[:, 0], ent[:, 1], ent[:, 2]
    L = np.float32(bl)
    shift = np.array([-L, np.float32(0), L], np.float32)
    img = x[j] + np.stack([shift[code // 9], shift[(code % 9) // 3], shift[code % 3]], 1)
    img[code == 13] = x[j[code == 13]]
    d = img - x[i]
    rsq = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    return int(np.count_nonzero(rsq < np.float32(rc) * np.float32(rc)))


def mfma_busy_cycles(kernel):
    """SQ_VALU_MFMA_BUSY_CYCLES per launch of the kernel (summed over the SIMDs) from the
    committed MFMA counter summary (the file states the command), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "pmc_mfma_water24k_mlp.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 9 and cells[0] == f"`{kernel}`":
            try:
                return float(cells[2]) * 1e6
            except ValueError:
                return None
    return None


def valu_busy(stage):
    """VALU busy share of the stage's kernel from the committed VALU counter summary (one
    rocprofv3 pass of this workload; the file states the command and the formula), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "pmc_valu_water24k_nnp_step.md")
    name = STAGE_KERNEL.get(stage)
    if name is None or not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 9 and cells[0] == f"`{name}`":
            try:
                return float(cells[4].rstrip(" %")) / 100.0
            except ValueError:
                return None
    return None


def cpu_baseline_leg(sysm, par, norm, x_end, drift, do_cpu, do_err):
    """The host side of the run: the reference's own CPU step and the oracle's restatement of the
    NNP step timed on the host cores, and the oracle as the checker of the GPU forces (FP64
    reference forces on the starting snapshot and on the liquid after the NVE run)."""
    cpu = errs = None
    if do_err:
        errs = {"start": force_errors(sysm, par, norm)}
        if x_end is not None:
            # the liquid after the NVE run of the timed form (positions as the engine holds them)
            snap = sysm.copy()
            snap.x = np.ascontiguousarray(x_end, np.float32)
            errs[f"after_{drift[HEAD]['ps']:g}ps_{HEAD}_nve"] = force_errors(snap, par, norm)
    if do_cpu:
        ref = cpu_reference(sysm)
        port = cpu_port(sysm, par, norm)
        if ref is not None:
            cpu = dict(ref)
            cpu["port_nnp_step"] = port
        else:
            cpu = port
    return cpu, errs


def refinement(steps, target_ps, iters):
    """The butane refinement workload (4000 chains, 16 000 sites, Ryckaert-Bellemans torsion):
    the target S(q) from an MD run with LJ epsilon x 1.1 (target_ps of FP64-Verlet MD, classical
    terms), then `iters` refinement iterations of `steps` MD steps each with the BF16 NNP on,
    g(r) + S(q) + RMDF every 10 steps and the pair table rewritten every 100 steps (each
    iteration continues from the previous one's positions and refined table), each timed."""
    from nnpengine_amd import workloads as W
    c = W.BUTANE_C3
    t0 = time.perf_counter()
    sysm = W.butane_c3_system()
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    target_steps = int(round(target_ps * 1000 / c["dt"]))
    target_steps -= target_steps % c["rebuild_every"]
    s_exp, _ = W.butane_target(sysm, target_steps)
    t_target = time.perf_counter() - t0
    log(f"[bench] refinement target: {target_steps} steps in {t_target:.1f} s")
    ft, ut = W.butane_tables(sysm)
    e = W.classical_engine(sysm, ft, ut, nnp=True)
    W.nnp_attach(e, sysm, HEAD)
    e.rmdf_init(W.rmdf_data(s_exp, sysm.ntype, c["q_max"]), 1.0, 0.5, target=0)
    kw = dict(dt=c["dt"], fp64=True, rebuild_every=c["rebuild_every"], pair=True, bonded=True, nnp=2,
              gr_every=c["gr_every"], gr_long_every=c["gr_every"], sq=True, rmdf=True, log_every=100)
    e.md_run(200, **kw)  # warm-up: graphs captured, the refinement under way
    e.sync()
    from nnpengine_amd import _lib
    k = _lib.TERMS.index("rmdf")
    it_s, res = [], []
    for it in range(iters):
        t0 = time.perf_counter()
        lg = e.md_run(steps, **kw)
        e.sync()
        it_s.append(time.perf_counter() - t0)
        if len(lg):
            res.append((float(lg[0, 2 * k]), float(lg[-1, 2 * k])))
    st = e.md_stats()
    e.close()
    t = float(sum(it_s))
    return {"workload": f"butane refinement: {sysm.natom} sites ({sysm.nmol} chains), L={sysm.bl:.1f} A, RB torsion, "
                        f"BF16 NNP ({HEAD}), g(r) + S(q) + RMDF every 10 steps, table rewrite every 100 steps",
            "iterations": iters, "steps_per_iteration": steps, "seconds": t,
            "iteration_seconds_min_max": [min(it_s), max(it_s)] if it_s else None,
            "ms_per_step": t / max(1, iters * steps) * 1e3, "atom_steps_per_s": sysm.natom * iters * steps / t if t else None,
            "target_ps": target_steps * c["dt"] / 1000, "target_seconds": t_target, "build_seconds": t_build,
            "rmdf_residual_first_iteration": list(res[0]) if res else None,
            "rmdf_residual_last_iteration": list(res[-1]) if res else None,
            "list_builds": int(st["nbuild"]), "skin_violations": int(st["skin_violations"])}


def force_errors(sysm, par, norm):
    """NNP and total force relative RMS vs the FP64 oracle on one snapshot."""
    from nnpengine_amd import model as M
    from nnpengine_amd import system as S
    from oracle import pyoracle
    from nnpengine_amd.engine import Engine
    ft, ut = S.lj_table(sysm, RC_PAIR, 900, S.WATER_LJ)
    sfp = M.sf_params(RC_NNP)
    o = pyoracle.Oracle(sysm, 1)
    o.neigh_init(RC_PAIR, SKIN, R_GR)
    # classical terms on the reference's half list, NNP terms on the full list
    o.rebuild(False)
    o.pair_set(RC_PAIR, ft, ut)
    o.clear_f()
    o.bonded_compute()
    o.pair_compute()
    f_cls64 = o.get("f")
    o.rebuild(True)
    o.sf_set(sfp["rc"], sfp["mu"], sfp["eta_r"], sfp["eta_a"])
    o.mlp_set(WIDTHS, par, 0.5, *norm)
    o.clear_f()
    G = o.sf_descriptor()
    _, _, dG = o.mlp_eval(G)
    o.sf_force(dG)
    f_nnp64 = o.get("f")
    f_tot64 = f_cls64 + f_nnp64

    def rr(a, b):
        return float(np.sqrt(((a - b) ** 2).mean() / (b ** 2).mean()))

    out = {}
    for name in ("fp32", "bf16", "fp16", "bf16x1"):
        e = make_engine(sysm, par, norm, name)
        e.build()
        e.sf_descriptor()
        e.clear_force()
        e.mlp_eval(False)
        e.sf_force(False)
        f_nnp = e.get("f")
        e.pair_compute(False)
        e.bonded_compute(False)
        e.sf_descriptor()
        e.mlp_eval(False)
        e.sf_force(False)
        f_tot = e.get("f")
        out[name] = {"nnp_force_rel_rms": rr(f_nnp, f_nnp64), "total_force_rel_rms": rr(f_tot, f_tot64)}
        e.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-err", action="store_true")
    ap.add_argument("--lat-samples", type=int, default=200, help="single-step replays for the latency percentiles")
    ap.add_argument("--stage-reps", type=int, default=20, help="launches per stage in the stage profile")
    ap.add_argument("--serve-requests", type=int, default=256, help="requests of the serving run (0: skip)")
    ap.add_argument("--serve-steps", type=int, default=500)
    ap.add_argument("--no-sqd", action="store_true")
    ap.add_argument("--drift-ps", type=float, default=20.0, help="NVE run per precision form for the drift (0: skip)")
    ap.add_argument("--refine-steps", type=int, default=2000, help="steps per butane refinement iteration (0: skip)")
    ap.add_argument("--refine-iters", type=int, default=50, help="refinement iterations")
    ap.add_argument("--refine-target-ps", type=float, default=200.0, help="MD run that makes the refinement target S(q)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); a sharded run is not "
              "part of this program", file=sys.stderr)
        sys.exit(2)
    from nnpengine_amd import model as M
    from nnpengine_amd import system as S

    sysm = S.water_box(NMOL, BL, seed=7)
    # network: seeded weights, descriptor standardisation from the starting snapshot's descriptors
    norm = gpu_standardisation(sysm)
    par = M.mlp_init(WIDTHS, sysm.ntype, 2024)
    log(f"[bench] system {sysm.natom} atoms")

    res = {}
    engines = {}
    for prec in ("bf16", "fp32", "bf16x1"):
        e = make_engine(sysm, par, norm, prec)
        t = time_steps(e, args.steps, args.warmup)
        res[prec] = t
        engines[prec] = e
        st = e.md_stats()
        log(f"[bench] {prec}: {args.steps} steps in {t:.3f} s = {t / args.steps * 1e3:.4f} ms/step; stats {st}")

    # per-step latency distribution of the BF16 form (events around single graph replays)
    eb = engines[HEAD]
    lat = []
    for _ in range(args.lat_samples):
        t0 = time.perf_counter()
        eb.md_run(1, log_every=0, **MD_KW)
        lat.append((time.perf_counter() - t0) * 1e3)
    lat = np.array(lat)
    # list sizes of the step as run (before the stage profile, whose Verlet stage moves the atoms
    # without new forces): entries of the pair and short lists, pairs within the SF cutoff
    n_long, n_ghost = eb.list_size(0)
    n_short, _ = eb.list_size(1)
    n_act = active_pairs(eb, sysm.bl, RC_NNP)

    # per-stage profile (HIP events on the engine stream) of the BF16 form
    stages = {}
    for st in ("pair", "bonded", "sf_desc", "mlp", "sf_adjoint", "sf_force", "rebuild", "gr_full", "rebuild_gr", "verlet", "sq",
               "copy", "mfma"):
        stages[st] = eb.profile(st, reps=args.stage_reps)
    # the peaks measured on this GPU beside the published ones: a 64 MiB streaming copy (read +
    # write) and an MFMA issue-rate probe (register operands, bf16 32x32x16)
    measured = {"hbm_copy_gbs": 2 * COPY_PROBE_BYTES / (stages["copy"] * 1e-3) / 1e9,
                "mfma_bf16_tflops": MFMA_PROBE_FLOP / (stages["mfma"] * 1e-3) / 1e12}
    measured["hbm_copy_frac_of_published"] = measured["hbm_copy_gbs"] / HBM_PEAK_GBS
    measured["mfma_bf16_frac_of_published"] = measured["mfma_bf16_tflops"] / MFMA_PEAK["bf16"]
    log("[bench] measured peaks:", measured)
    ms32_mlp = engines["fp32"].profile("mlp", reps=args.stage_reps)
    ms_split_mlp = engines["bf16"].profile("mlp", reps=args.stage_reps)
    log("[bench] stage ms:", {k: round(v, 4) for k, v in stages.items()}, "fp32 mlp", round(ms32_mlp, 4))

    # roofline of the dominant kernel: the largest share of the step (stage time x launches per
    # step; rebuild and the full g(r) tally run every 10th step)
    N = sysm.natom
    flop_atom = 4 * sum(WIDTHS[i] * WIDTHS[i + 1] for i in range(len(WIDTHS) - 1))
    # (the full g(r) tally runs on the rebuild steps, folded into the rebuild: stage rebuild_gr)
    # (the verlet stage times both half-steps; in the NNP step the second rides on the SF force
    # pass, so about half of it is a launch of its own)
    per_step = {"pair": 1.0, "bonded": 1.0, "sf_desc": 1.0, "mlp": 1.0, "sf_adjoint": 1.0, "sf_force": 1.0,
                "verlet": 0.5, "rebuild_gr": 0.1, "sq": 0.1}
    share = {k: stages[k] * f for k, f in per_step.items()}
    # the dominant KERNEL: rebuild_gr and sq are several launches each (the rebuild's largest,
    # k_build_cells, is well under half of its stage), so the pick is among one-kernel stages
    one_kernel = ("pair", "bonded", "sf_desc", "mlp", "sf_adjoint", "sf_force", "verlet")
    dom = max(one_kernel, key=share.get)
    nmt = sysm.ntype * 4 * 11 + sysm.ntype * 16   # moments per atom
    nadj = sysm.ntype * 64                        # adjoint row: one 64-float slice per element
    D = WIDTHS[0]
    # algorithmic HBM bytes per launch: each input read once, each output written once
    # (list entries 4 B; xq 16 B, typ 4 B, f 16 B per atom; moments/adjoints 4*nmt B; G/dG 4*D B)
    alg_bytes = {"pair": 4 * n_long + N * (16 + 4 + 16),
                 "sf_desc": 4 * n_short + N * (16 + 4 + 4 * nmt + 4 * D),
                 "sf_adjoint": N * 4 * (nmt + D + nadj),
                 # (the active lists; own slices FP32 + the FP16 copy the 16-bit forms gather from)
                 "sf_force": 4 * n_act + N * (16 + 4 + 4 * nadj + 2 * nadj + 16 + 16),
                 "bonded": N * (16 + 16 + 16),
                 "verlet": N * 164,
                 "rebuild": 4 * (n_long + n_short) + N * 2 * (16 + 16 + 16 + 4 + 4 + 16 + 3 * 32),
                 "rebuild_gr": 4 * (n_long + n_short) + N * 2 * (16 + 16 + 16 + 4 + 4 + 16 + 3 * 32),
                 "gr_full": N * (16 + 4),
                 "sq": 8 * 1000 * 600}
    if dom == "mlp":
        roof = {"kernel": "k_mlp<bf16>", "bound": "mfma", "achieved": flop_atom * N / (stages["mlp"] * 1e-3) / 1e12,
                "peak": MFMA_PEAK["bf16"], "unit": "TFLOP/s"}
    else:
        roof = {"kernel": dom, "bound": "hbm", "achieved": alg_bytes[dom] / (stages[dom] * 1e-3) / 1e9,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "algorithmic_bytes": alg_bytes[dom]}
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["traffic"] = hbm_traffic(dom)
    # the symmetry-function passes against the vector-ALU roofline: algorithmic FP32 operations
    # per directed pair within the cutoff (FMA = 2, exp / rsqrt = 1; counted in the design notes)
    # over the non-packed FP32 rate
    alg_flops = {"sf_force": SF_FORCE_FLOP_PAIR * n_act, "sf_desc": SF_DESC_FLOP_PAIR * n_act}
    if dom in alg_flops:
        ach = alg_flops[dom] / (stages[dom] * 1e-3) / 1e12
        roof["valu"] = {"algorithmic_flop": alg_flops[dom], "flop_per_pair": alg_flops[dom] / n_act,
                        "active_pairs": n_act, "achieved_tflops": ach, "peak_tflops": VALU_PEAK_TFLOPS,
                        "frac_valu": ach / VALU_PEAK_TFLOPS}
    # the SF and pair kernels are bound by their vector-ALU issue and gather latency rather than
    # by HBM: the committed VALU counter pass gives the kernel's VALU busy share beside frac
    roof["valu_busy"] = valu_busy(dom)
    roof["step_share"] = {k: round(v / sum(share.values()), 4) for k, v in share.items()}
    mlp_roof = {"achieved_tflops_bf16": flop_atom * N / (stages["mlp"] * 1e-3) / 1e12,
                "achieved_tflops_fp32": flop_atom * N / (ms32_mlp * 1e-3) / 1e12,
                "flop_per_atom": flop_atom}
    mlp_roof["frac_mfma_bf16"] = mlp_roof["achieved_tflops_bf16"] / MFMA_PEAK["bf16"]
    mlp_roof["frac_mfma_fp32"] = mlp_roof["achieved_tflops_fp32"] / MFMA_PEAK["fp32"]
    # MFMA busy share by counter: the per-launch SQ_VALU_MFMA_BUSY_CYCLES of the committed MFMA
    # pass over the SIMD cycles of this run's stage time
    busy = {"bf16": mfma_busy_cycles(STAGE_KERNEL["mlp"]), "fp32": mfma_busy_cycles("k_mlp<0, 3>")}
    mlp_roof["mfma_busy_share"] = {
        "bf16": busy["bf16"] / (1024 * stages["mlp"] * 1e-3 * CLOCK_HZ) if busy["bf16"] else None,
        "fp32": busy["fp32"] / (1024 * ms32_mlp * 1e-3 * CLOCK_HZ) if busy["fp32"] else None}
    # every stage against the HBM roofline (algorithmic bytes / live stage time), with the HBM
    # traffic and VALU busy share of its kernel from the committed counter summaries where the
    # stage is one kernel (the rebuild is a dozen launches: its time and bytes cover them all)
    stage_roof = {}
    for st in ("rebuild", "pair", "bonded", "sf_desc", "sf_adjoint", "sf_force", "verlet"):
        ach = alg_bytes[st] / (stages[st] * 1e-3) / 1e9
        stage_roof[st] = {"algorithmic_bytes": alg_bytes[st], "achieved_gbs": ach, "frac_hbm": ach / HBM_PEAK_GBS,
                          "traffic": hbm_traffic(st), "valu_busy": valu_busy(st)}
        if st in alg_flops:
            stage_roof[st]["frac_valu"] = alg_flops[st] / (stages[st] * 1e-3) / 1e12 / VALU_PEAK_TFLOPS

    sqd = None if args.no_sqd else sqd_timing(eb)

    for e_ in engines.values():
        e_.close()
    serve = None
    if args.serve_requests > 0:
        serve = serving(par, norm, args.serve_requests, args.serve_steps)
        log("[bench] serving:", {k: v for k, v in serve.items() if k != "latency_s"}, serve["latency_s"])
    drift, x_end = {}, None
    if args.drift_ps > 0:
        for prec in ("bf16", "fp16", "fp32", "bf16x1"):
            drift[prec], xe = nve_drift(sysm, par, norm, prec, args.drift_ps)
            if prec == HEAD:
                x_end = xe
            log(f"[bench] NVE {prec}: {drift[prec]}")
    refine = None
    if args.refine_steps > 0:
        refine = refinement(args.refine_steps, args.refine_target_ps, args.refine_iters)
        log("[bench] refinement:", refine)
    cpu, errs = cpu_baseline_leg(sysm, par, norm, x_end, drift, not args.no_cpu, not args.no_err)

    t_b, t_f = res[HEAD], res["fp32"]
    value = N * args.steps / t_b
    out = {
        "metric": "NNP-MD atom-steps/s @24k atoms, BF16-MFMA vs FP32 at stated err; p99 latency",
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_b / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "precision_form": "NNP hidden layers in BF16 on MFMA (one BF16 term per value, FP32 accumulation, "
                          "FP32 bias/tanh/epilogues); descriptors, tabulated pair terms and the own-atom "
                          "adjoints FP32; the neighbours' adjoints (dE/dmoment) gathered in FP16 by the SF "
                          "force pass; Verlet, energy/virial sums and S(q) FP64",
        "data": "synthetic (seeded water lattice, Maxwell-Boltzmann 300 K, random-init NNP weights)",
        "config": {"workload": "24000-atom water NNP-MD step (8000 molecules, L=62.1 A, LJ+Coulomb table 9 A, "
                               "symmetry-function NNP 6 A D=80, MLP 3x128 tanh, lists 11 A rebuilt every 10 steps, "
                               "g(r) 0-12 A + S(q) every 10 steps, FP64 Verlet dt=0.5 fs)",
                   "natom": N, "dt_fs": DT},
        "ns_per_day": args.steps / t_b * DT * 0.0864 / 1.0,
        "fp32_form": {"value": N * args.steps / t_f, "ms_per_step": t_f / args.steps * 1e3},
        "bf16_speedup_vs_fp32": t_f / t_b,
        # the BF16 network with weights (and backward gradients) as hi + lo BF16 pairs: ten times
        # smaller NNP force error, slower network
        "bf16_split_form": {"value": N * args.steps / res["bf16"], "ms_per_step": res["bf16"] / args.steps * 1e3,
                            "speedup_vs_fp32": t_f / res["bf16"], "mlp_ms": ms_split_mlp},
        "force_error_vs_fp64": errs,
        "nve_drift": drift or None,
        "step_latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "serving": serve,
        "stage_ms": stages,
        "mlp": mlp_roof,
        "roofline": roof,
        "stage_roofline": stage_roof,
        "measured_peaks": measured,
        "sqd": sqd,
        "refinement": refine,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
