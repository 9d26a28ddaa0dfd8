"""Time-step throughput of the granular compaction workload on one MI355X.

Workload: 112,000 spheres (radii uniform in [0.04, 0.06] m, seed 2002)
deposited in a 5 x 5 m open box, mu = 0.3, h = 1e-3 s, APGD tolerance 1e-4
with warm starting (about 256k contacts, D of 768k x 336k).  The deposit is
not an equilibrium of the contact problem: its first ~20 steps need 10^4
APGD iterations each, so the pile is first settled for 30 untimed steps and
every solver is then timed from that same settled state.  One step is
System::DoTimeStep (System.cu:544-602): collision detection, assembly, APGD
to the tolerance, update.  A separate leg steps the pile under a free lid as
heavy as the spheres (the reference's compaction driver) to show what a
pressed pile costs.

The headline `value` is time-steps/s of the mixed-precision solver (FP32
D, M^-1 and products, FP64 iterates, tests and update).  The same run times
the FP64 graph-driven solver and the plain FP64 solver (one kernel per
vector operation, host-driven like the reference's CUSP calls) from the same
initial state, and the CPU oracle on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-steps S] [--no-cpu]
"""
import argparse
import ctypes
import copy
import json
import os
import sys
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, _HERE)

HBM_PEAK_GBS = 8000.0
TOL = 1e-4


def _hip():
    lib = ctypes.CDLL("libamdhip64.so")
    lib.hipDeviceSynchronize.restype = ctypes.c_int
    return lib


def p1_bytes_per_contact(mixed: bool, fresh: bool = True) -> int:
    """Algorithmic bytes the contact pass P1 moves per contact (3 rows).

    Reads gamma_k, gamma_{k-1} (FP64), N gamma_{k-1} and r (TM), the stored
    frame (FP64 n, or FP32 n and u); writes gamma_new (FP64) and the
    world-frame impulse J = F^T gamma_new (3 TM).  N gamma_k: when the
    previous pass accepted gamma_k (fresh), P1 forms it from the contact's two
    body ids and the bodies' M^-1 D^T gamma_k (3 TM each) and stores it (3
    TM); otherwise it reads it back (3 TM).  J's second copy in the B-side
    order (and the index it is stored at) is a layout choice and is not
    counted.
    """
    s = 4 if mixed else 8
    n_gamma_k = (8 + 6 * s + 3 * s) if fresh else 3 * s
    return 2 * 24 + 6 * s + 24 + n_gamma_k + 24 + 3 * s


def p2_bytes(nc: int, nb: int, mixed: bool) -> int:
    """Algorithmic bytes of the body pass P2: each contact's J read by both
    of its bodies (3 TM each time); per body its three index offsets, M^-1,
    the store of M^-1 D^T gamma_new (3 TM) and, in the mixed solver, the
    D^T gamma ring (two reads and one store of 3 TM) for the Lipschitz term."""
    s = 4 if mixed else 8
    return nc * 6 * s + nb * (12 + s + 3 * s + (9 * s if mixed else 0))


SETTLE_STEPS = 30


def settled_scene(settle_steps=SETTLE_STEPS):
    """The compaction pile after `settle_steps` untimed mixed-precision steps
    (iteration cap 20000): positions and velocities replace the deposit's."""
    from dvide_amd import scenes, system

    sc = scenes.compaction(2002, lid=scenes.LID_NONE)
    g = system.from_scene(sc, system.BUILD_MIXED, tol=TOL, max_iterations=20000)
    its = []
    for _ in range(settle_steps):
        g.step()
        its.append(g.stats().iterations)
    sc.pos = g.p.reshape(-1, 3).copy()
    sc.vel = g.v.reshape(-1, 3).copy()
    g.close()
    return sc, its


def run_gpu(scene, solver, steps, warmup, profile_kernel=0, max_iterations=None):
    from dvide_amd import system

    g = system.from_scene(scene, solver, tol=TOL, max_iterations=max_iterations)
    hip = _hip()
    try:
        for _ in range(warmup):
            g.step()
        g.setProfileKernel(profile_kernel)
        hip.hipDeviceSynchronize()
        lat, its, res, nc, passes, solve, kms, kcnt, kstep = [], [], [], [], [], [], 0.0, 0, []
        any_tuning = False
        t0 = time.perf_counter()
        for _ in range(steps):
            ts = time.perf_counter()
            g.step()
            lat.append(time.perf_counter() - ts)
            st = g.stats()
            its.append(st.iterations)
            res.append(st.residual)
            nc.append(st.num_contacts)
            passes.append(st.passes)
            solve.append(st.ms_solve)
            kms += st.prof_kernel_ms
            kcnt += st.prof_kernel_count
            kstep.append(round(1e3 * st.prof_kernel_ms / st.prof_kernel_count, 3) if st.prof_kernel_count else None)
            any_tuning = any_tuning or bool(st.schedule_tuning)
        hip.hipDeviceSynchronize()
        elapsed = time.perf_counter() - t0
        last = g.stats()
        return dict(
            elapsed=elapsed, lat=np.array(lat), iterations=np.array(its), residual=np.array(res),
            contacts=np.array(nc), passes=np.array(passes), solve_ms=np.array(solve), kernel_ms=kms,
            kernel_count=kcnt, kernel_us_steps=kstep,
            phases=dict(collision=last.ms_collision, assembly=last.ms_assembly, solve=last.ms_solve,
                        update=last.ms_update),
            nb=g.numBodies, v=g.v, p=g.p,
            schedule=dict(rows_per_thread=last.rows_per_thread, xcd_order=last.xcd_order,
                          measured_in_timed_steps=int(any_tuning)),
        )
    finally:
        g.close()


def copy_calibration(nbytes=1 << 30):
    """A plain copy kernel of known size (torch elementwise copy, nbytes read
    and nbytes written), run under rocprofv3 --pmc to scale FETCH_SIZE and
    WRITE_SIZE into HBM bytes."""
    import torch

    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.add(x, 0.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": nbytes}), flush=True)


def _short_kernel_name(name):
    name = name.replace("dv::(anonymous namespace)::", "").replace("void ", "")
    return name.split("(")[0]


def summarize_rocprof(path):
    """Markdown table of a rocprofv3 --output-format csv file: per kernel, the
    launch count and durations (kernel_stats.csv), or the median value of
    each counter over the launches (counter_collection.csv)."""
    import csv

    rows = list(csv.DictReader(open(path)))
    if rows and "AverageNs" in rows[0]:
        print("| kernel | launches | avg us | min us | max us | total ms |")
        print("|---|---|---|---|---|---|")
        for r in rows[:16]:
            print(f"| `{_short_kernel_name(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
                  f"{float(r['MinNs']) / 1e3:.2f} | {float(r['MaxNs']) / 1e3:.2f} | "
                  f"{float(r['TotalDurationNs']) / 1e6:.2f} |")
        return
    vals = {}
    for r in rows:
        key = (_short_kernel_name(r["Kernel_Name"]), r["Counter_Name"])
        vals.setdefault(key, []).append(float(r["Counter_Value"]))
    print("| kernel | counter | launches | median | mean |")
    print("|---|---|---|---|---|")
    for (k, c), v in sorted(vals.items(), key=lambda kv: -len(kv[1])):
        print(f"| `{k}` | {c} | {len(v)} | {float(np.median(v)):.0f} | {float(np.mean(v)):.0f} |")


def lid_pressed(steps=5):
    """The deposit under a free lid as heavy as the spheres (the reference's
    compaction driver): per-step APGD iterations and times of both GPU
    solvers from the first step, iteration cap 20000."""
    from dvide_amd import scenes, system

    sc = scenes.compaction(2002)
    out = {}
    for name, solver in (("mixed", system.BUILD_MIXED), ("fp64", system.BUILD_FP64)):
        g = system.from_scene(sc, solver, tol=TOL, max_iterations=20000)
        its, res, ms = [], [], []
        for _ in range(steps):
            g.step()
            st = g.stats()
            its.append(st.iterations)
            res.append(float(f"{st.residual:.3g}"))
            ms.append(round(st.ms_total, 1))
        out[name] = {"iterations": its, "residual": res, "ms": ms}
        g.close()
    return out


def shell_bed_leg(settle=400, steps=20):
    """Spheres on a fixed ANCF shell: 19,200 spheres r ~ U[0.02, 0.03] in
    eight jittered lattice layers over a fixed 20 x 80 mesh of 0.1 m elements
    (1,701 nodes; 8 contact points per element side, 57,600 points), mu 0.4,
    h 5e-4, APGD tol 1e-4, warm start.  `settle` untimed steps let the layers
    land, then `steps` timed steps per graph-driven solver: time-steps/s,
    contacts, iterations and residuals, and the vertical contact force the
    shell's nodes carry as a share of the spheres' weight."""
    from dvide_amd import scenes, system

    sc = scenes.shell_bed(3003, nx=20, ny=80, per_side=(24, 100), layers=8, tol=TOL)
    weight = float(sc.mass.sum()) * 9.81
    w = sc.shell.points - 2
    out = {"spheres": sc.n, "shell_nodes": sc.shell.n_nodes, "contact_points": sc.shell.n_elements * w * w,
           "settle_steps": settle}
    for name, build in (("mixed", system.BUILD_MIXED), ("fp64", system.BUILD_FP64)):
        g = system.from_scene(sc, build, tol=TOL, max_iterations=REFERENCE_CAP)
        for _ in range(settle):
            g.step()
        nc, its, res = [], [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            g.step()
            st = g.stats()
            nc.append(st.num_contacts)
            its.append(st.iterations)
            res.append(st.residual)
        el = time.perf_counter() - t0
        fc = g.f_contact
        out[name] = {"time_steps_per_s": round(steps / el, 2), "steps": steps, "contacts_mean": int(np.mean(nc)),
                     "iterations_mean": round(float(np.mean(its)), 1), "iterations_max": int(np.max(its)),
                     "residual_max": float(np.max(res)), "steps_converged": int(sum(r < TOL for r in res)),
                     "shell_carries_weight_share": round(float(-fc[3 * sc.n:].reshape(-1, 9)[:, 1].sum()) / weight,
                                                         4)}
        g.close()
    return out


def fp64_leg(scene, steps, warmup, cap):
    """The FP64 graph-driven solver alone on the compaction pile: time-steps/s,
    its P1 launches burst-timed over the timed steps, its P2 launches over two
    more steps from the same state, and P1's algorithmic bytes against HBM peak
    (248 B per contact after an accepted iterate, 192 after a backtrack,
    weighted by the share of fresh passes)."""
    from dvide_amd import system

    r = run_gpu(scene, system.BUILD_FP64, steps, warmup, profile_kernel=system.PROFILE_P1, max_iterations=cap)
    p2 = run_gpu(scene, system.BUILD_FP64, 2, 0, profile_kernel=system.PROFILE_P2, max_iterations=cap)
    nc = int(np.mean(r["contacts"]))
    p1_us = 1e3 * r["kernel_ms"] / max(r["kernel_count"], 1)
    p2_us = 1e3 * p2["kernel_ms"] / max(p2["kernel_count"], 1)
    fresh = float(np.sum(r["iterations"])) / max(float(np.sum(r["passes"])), 1.0)
    b = fresh * p1_bytes_per_contact(False, True) + (1 - fresh) * p1_bytes_per_contact(False, False)
    gbs = b * nc / (p1_us * 1e-6) / 1e9 if p1_us > 0 else 0.0
    return {"time_steps_per_s": round(steps / r["elapsed"], 3), "steps": steps, "contacts_mean": nc,
            "bodies": int(r["nb"]), "iterations_mean": round(float(np.mean(r["iterations"])), 1),
            "passes_mean": round(float(np.mean(r["passes"])), 1), "fresh_share": round(fresh, 3),
            "P1_us": round(p1_us, 3), "P2_us": round(p2_us, 3),
            "p1_bytes_per_contact": round(b, 1), "p1_achieved_GBs": round(gbs, 1),
            "p1_frac_of_8TBs": round(gbs / 8000.0, 4), "schedule": r["schedule"]}


def _gen_many(specs):
    """Scenes generated on the host cores in parallel (the generator releases
    the interpreter lock)."""
    from concurrent.futures import ThreadPoolExecutor

    from dvide_amd import scenes

    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    with ThreadPoolExecutor(workers) as ex:
        return list(ex.map(lambda a: scenes.pile(*a), specs))


# iteration caps of the batched and served legs: the deposit transient of a
# fresh pile is settled under SETTLE_CAP, and every timed solve runs to the
# tolerance under the reference's default cap (APGD.cu:11)
SETTLE_CAP = 20000
REFERENCE_CAP = 100000
SERVE_SETTLE = 6
CAPACITY_WINDOW_S = 8.0


def batched(n_scenes=512, steps=2, settle=3, seed=4004):
    """512 independent piles, contacts log-uniform in [1k, 100k], radii
    U[0.03, 0.06] m, mu U[0.1, 0.5]; one mixed-precision APGD solve of all of
    them per step in the same graph launches.  The fresh deposits are first
    settled for `settle` untimed batch steps; the timed steps solve every
    scene to the tolerance under the reference's iteration cap, and the
    converged count and the largest residual are reported beside the rate."""
    from dvide_amd import system

    rng = np.random.default_rng(seed)
    targets = np.exp(rng.uniform(np.log(1e3), np.log(1e5), n_scenes)).astype(int)
    mus = rng.uniform(0.1, 0.5, n_scenes)
    t0 = time.perf_counter()
    scs = _gen_many([(seed + i, int(targets[i]), 0.03, 0.06, float(mus[i])) for i in range(n_scenes)])
    gen = time.perf_counter() - t0
    systems = [system.from_scene(sc, system.BUILD_MIXED, tol=TOL, max_iterations=SETTLE_CAP) for sc in scs]
    batch = system.Batch(system.BUILD_MIXED)
    for g in systems:
        batch.add(g)
    t0 = time.perf_counter()
    for _ in range(settle):
        batch.step()
    settle_s = time.perf_counter() - t0
    for g in systems:
        g.solver.maxIterations = REFERENCE_CAP
    hip = _hip()
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    stats, res = [], []
    for _ in range(steps):
        batch.step()
        stats.append(batch.stats())
        res.append([g.stats().residual for g in systems if g.stats().num_contacts])
    hip.hipDeviceSynchronize()
    dt = time.perf_counter() - t0
    contacts = np.mean([s.total_contacts for s in stats])
    res_all = np.concatenate([np.asarray(r) for r in res])
    out = {
        "scenes": n_scenes,
        "scene_steps_per_s": round(n_scenes * steps / dt, 2),
        "contact_steps_per_s": round(contacts * steps / dt, 1),
        "ms_per_batch_step": round(1e3 * dt / steps, 2),
        "contacts_per_step": int(contacts),
        "scene_steps_solved": int(res_all.size),
        "scene_steps_converged": int(np.sum(res_all < TOL)),
        "residual_max": float(res_all.max()),
        "iteration_cap": REFERENCE_CAP,
        "max_iterations": int(max(s.max_iterations for s in stats)),
        "mean_iterations": round(float(np.mean([s.total_iterations / n_scenes for s in stats])), 1),
        "ms_assembly": round(float(np.mean([s.ms_assembly for s in stats])), 2),
        "ms_solve": round(float(np.mean([s.ms_solve for s in stats])), 2),
        "schedule": {"rows_per_thread": stats[-1].rows_per_thread, "xcd_order": stats[-1].xcd_order,
                     "wg_per_cu_contacts": stats[-1].wg_per_cu_contacts,
                     "wg_per_cu_bodies": stats[-1].wg_per_cu_bodies,
                     "measured_in_timed_steps": int(any(x.schedule_tuning for x in stats))},
        "settle_batch_steps": settle,
        "settle_s": round(settle_s, 1),
        "scene_generation_s": round(gen, 1),
    }
    batch.close()
    for g in systems:
        g.close()
    return out


def served(n_scenes=256, n_requests=512, load=0.7, seed=5005, max_slots=64):
    """Poisson stream of time-step requests over 256 resident scenes, sizes
    log-uniform in [1k, 200k] contacts, radii U[0.02, 0.03] m, mu 0.4; the
    server admits by free slot and HBM budget and overlaps the solves.  The
    rate is `load` times the throughput measured in a closed-loop warm-up."""
    from dvide_amd import system

    rng = np.random.default_rng(seed)
    targets = np.exp(rng.uniform(np.log(1e3), np.log(2e5), n_scenes)).astype(int)
    scs = _gen_many([(seed + i, int(targets[i]), 0.02, 0.03, 0.4) for i in range(n_scenes)])
    systems = [system.from_scene(sc, system.BUILD_MIXED, tol=TOL, max_iterations=SETTLE_CAP) for sc in scs]
    srv = system.Server(system.BUILD_MIXED, max_slots=max_slots)
    # the fresh deposits' transient: untimed requests, every scene in turn
    t0 = time.perf_counter()
    for _ in range(SERVE_SETTLE):
        for g in systems:
            srv.submit(g)
        srv.drain()
    settle_s = time.perf_counter() - t0
    for g in systems:
        g.solver.maxIterations = REFERENCE_CAP
    # closed loop: every scene in flight, each resubmitted as soon as its
    # request completes; the completions in a fixed window give the sustained
    # capacity (a pass over every scene once would be timed by its slowest
    # request instead)
    scene_of = {srv.submit(g): i for i, g in enumerate(systems)}
    t0 = time.perf_counter()
    completed = 0
    while time.perf_counter() - t0 < CAPACITY_WINDOW_S:
        for tk in srv.poll():
            completed += 1
            i = scene_of.pop(tk)
            scene_of[srv.submit(systems[i])] = i
    capacity = completed / (time.perf_counter() - t0)
    srv.drain()
    rate = load * capacity
    # open loop: Poisson arrivals at `rate`, scenes drawn uniformly; a request
    # for a scene with one in flight waits for it
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    which = rng.integers(0, n_scenes, n_requests)
    busy, queued, ticket_of = set(), {}, {}
    submitted_at, latency, residuals = {}, [], []
    start = time.perf_counter()
    nxt = 0
    while len(latency) < n_requests:
        now = time.perf_counter() - start
        while nxt < n_requests and arrivals[nxt] <= now:
            queued.setdefault(int(which[nxt]), []).append(arrivals[nxt])
            nxt += 1
        for sc_i in list(queued):
            if sc_i not in busy and queued[sc_i]:
                t_arr = queued[sc_i].pop(0)
                tk = srv.submit(systems[sc_i])
                ticket_of[tk] = sc_i
                submitted_at[tk] = t_arr
                busy.add(sc_i)
            if not queued[sc_i]:
                del queued[sc_i]
        if srv.pending():
            for tk in srv.poll():
                latency.append(time.perf_counter() - start - submitted_at.pop(tk))
                sc_i = ticket_of.pop(tk)
                busy.discard(sc_i)
                st = systems[sc_i].stats()
                if st.num_contacts:
                    residuals.append(st.residual)
        elif nxt < n_requests:
            time.sleep(max(0.0, min(0.001, arrivals[nxt] - now)))
    total = time.perf_counter() - start
    lat = np.array(latency) * 1e3
    srv.close()
    for g in systems:
        g.close()
    return {
        "scenes": n_scenes,
        "requests": n_requests,
        "offered_rate_per_s": round(rate, 2),
        "throughput_per_s": round(n_requests / total, 2),
        "closed_loop_capacity_per_s": round(capacity, 2),
        "capacity_window_s": CAPACITY_WINDOW_S,
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2),
                       "p99": round(float(np.percentile(lat, 99)), 2),
                       "mean": round(float(lat.mean()), 2)},
        "settle_requests_per_scene": SERVE_SETTLE,
        "settle_s": round(settle_s, 1),
        "requests_solved": len(residuals),
        "requests_converged": int(np.sum(np.asarray(residuals) < TOL)),
        "residual_max": float(max(residuals)) if residuals else None,
        "iteration_cap": REFERENCE_CAP,
        "max_slots": max_slots,
        "contacts_mean": int(np.mean(targets)),
    }


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def host_cores():
    """The host cores this process may use: the affinity mask, within
    OMP_NUM_THREADS when that is set."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(env)) if env.isdigit() and int(env) > 0 else n


def cpu_baseline(scene, steps, threads, warm_fill=True):
    """The oracle (FP64 C++/OpenMP restatement) on `threads` host cores,
    from the settled state: with `warm_fill`, one untimed step first fills
    its warm-start gamma; then `steps` timed steps.  Returns the figures and
    the oracle's state after its first step, which the error check below
    compares the GPU solvers with."""
    import oracle

    oracle.set_threads(threads)
    o = oracle.OracleSystem(scene, tol=TOL)
    first = None
    if warm_fill:
        o.step()
        first = dict(v=o.vec(oracle.V), f_contact=o.vec(oracle.F_CONTACT), gamma=o.vec(oracle.GAMMA),
                     iterations=o.iterations, residual=o.residual)
    its = []
    t0 = time.perf_counter()
    for _ in range(steps):
        o.step()
        its.append(o.iterations)
    dt = time.perf_counter() - t0
    start = "after one untimed step" if warm_fill else "its first step, gamma from zero"
    out = dict(value=steps / dt, unit="time-steps/s", cores=threads, kind="port",
               sample=f"{steps} time step(s) of the same workload from the settled state ({start}), "
                      f"APGD to tol {TOL} ({'/'.join(map(str, its))} iterations), FP64, "
                      f"{threads} thread(s), {dt:.1f} s")
    return out, first


def oracle_threads(n):
    import oracle

    oracle.set_threads(n)


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _one_step_errors(scene, oracle_first, builds, converged=None):
    """One step of each GPU build from `scene` against the oracle's step from
    the same state at the same tolerance (FP64): relative errors of the free
    bodies' velocities and contact forces, the gamma error, and the residual
    the oracle evaluates in FP64 on the GPU's gamma.  With `converged` (the
    oracle's step at a tight tolerance), also each solution's velocity error
    against it, the oracle's own at the same tolerance included: at a
    residual of 1e-4 two solvers can stop at different points, and the
    distance to the converged solution says how far each is."""
    import oracle
    from dvide_amd import system

    free = np.repeat(np.asarray(scene.fixed) == 0, 3)
    n3 = free.size  # the bodies' DOFs (a shell's follow them)
    oracle_first = dict(oracle_first, v=oracle_first["v"][:n3], f_contact=oracle_first["f_contact"][:n3])
    if converged is not None:
        converged = dict(converged, v=converged["v"][:n3])
    o_asm = oracle.OracleSystem(scene, tol=TOL)
    o_asm.assemble()
    out = {}
    for name, build in builds:
        g = system.from_scene(scene, build, tol=TOL, max_iterations=REFERENCE_CAP)
        g.step()
        st = g.stats()
        out[name] = {
            "contacts": st.num_contacts,
            "iterations": st.iterations,
            "residual": float(f"{st.residual:.3g}"),
            "residual_by_oracle": float(f"{o_asm.residual_of(g.gamma):.3g}"),
            "rel_v": float(f"{_rel(g.v[:n3][free], oracle_first['v'][free]):.3g}"),
            "rel_f_contact": float(f"{_rel(g.f_contact[:n3][free], oracle_first['f_contact'][free]):.3g}"),
            "rel_gamma": float(f"{_rel(g.gamma, oracle_first['gamma']):.3g}"),
        }
        if converged is not None:
            out[name]["rel_v_vs_converged"] = float(f"{_rel(g.v[:n3][free], converged['v'][free]):.3g}")
        g.close()
    if converged is not None:
        out["oracle_rel_v_vs_converged"] = float(f"{_rel(oracle_first['v'][free], converged['v'][free]):.3g}")
        out["converged_tol"] = converged["tol"]
        out["converged_residual"] = float(f"{converged['residual']:.3g}")
    return out


def _oracle_first_step(scene, tol=TOL, max_iterations=REFERENCE_CAP):
    import oracle

    o = oracle.OracleSystem(scene, tol=tol, max_iterations=max_iterations)
    o.step()
    return dict(v=o.vec(oracle.V), f_contact=o.vec(oracle.F_CONTACT), gamma=o.vec(oracle.GAMMA),
                iterations=o.iterations, residual=o.residual, tol=tol)


def error_check(headline_scene, headline_first):
    """The mixed solver against the FP64 oracle, per workload, at matched
    residual (both solved to tol 1e-4 from the same state); the FP64
    graph-driven build beside it shows how far two FP64 solutions at that
    tolerance lie apart.  gamma is unique only on the single-layer scene
    (one contact per sphere); on piles N is rank-deficient, so v and
    f_contact carry the comparison there."""
    import oracle
    from dvide_amd import scenes, system

    builds = (("mixed", system.BUILD_MIXED), ("fp64", system.BUILD_FP64))
    rep = {"tol": TOL}
    rep["compaction_112k"] = _one_step_errors(headline_scene, headline_first, builds)
    # the 1,000-sphere lattice drop, in contact after 100 steps of the oracle
    lat = scenes.lattice_box(1001)
    o = oracle.OracleSystem(lat)
    for _ in range(100):
        o.step()
    lat.pos = o.vec(oracle.P).reshape(-1, 3).copy()
    lat.vel = o.vec(oracle.V).reshape(-1, 3).copy()
    rep["lattice_1000_step_101"] = _one_step_errors(lat, _oracle_first_step(lat), builds,
                                                    _oracle_first_step(lat, tol=1e-7, max_iterations=30000))
    # four members of the 512-scene batch family (seed 4004 + i), first step
    rng = np.random.default_rng(4004)
    targets = np.exp(rng.uniform(np.log(1e3), np.log(1e5), 512)).astype(int)
    mus = rng.uniform(0.1, 0.5, 512)
    fam = {}
    for i in (0, 1, 2, 3):
        sc = scenes.pile(4004 + i, int(targets[i]), 0.03, 0.06, float(mus[i]))
        fam[f"member_{i}"] = _one_step_errors(sc, _oracle_first_step(sc), builds,
                                              _oracle_first_step(sc, tol=1e-6, max_iterations=8000))
    rep["batch_members"] = fam
    # spheres on a fixed shell (the shell_bed leg's pile once landed: 400 mixed
    # steps on the GPU), one step from that state
    sb = scenes.shell_bed(3003, nx=20, ny=80, per_side=(24, 100), layers=8, tol=TOL)
    g = system.from_scene(sb, system.BUILD_MIXED, tol=TOL, max_iterations=REFERENCE_CAP)
    for _ in range(400):
        g.step()
    sb.pos, sb.vel = g.p[:3 * sb.n].reshape(-1, 3).copy(), g.v[:3 * sb.n].reshape(-1, 3).copy()
    g.close()
    rep["shell_bed_19k"] = _one_step_errors(sb, _oracle_first_step(sb), builds,
                                            _oracle_first_step(sb, tol=1e-6, max_iterations=20000))
    # gamma on the single layer (unique), at 1e-4 and against the converged 1e-8 solution
    lay = scenes.layer(7)
    rep["layer_256_gamma_unique"] = _one_step_errors(lay, _oracle_first_step(lay), builds)
    tight = _oracle_first_step(lay, tol=1e-8)
    g = system.from_scene(lay, system.BUILD_MIXED, tol=TOL, max_iterations=REFERENCE_CAP)
    g.step()
    rep["layer_256_gamma_unique"]["mixed"]["rel_gamma_vs_converged"] = float(
        f"{_rel(g.gamma, tight['gamma']):.3g}")
    g.close()
    return rep


def progress(msg):
    """A progress line on stderr (the result line alone goes to stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--settle", type=int, default=SETTLE_STEPS, help="untimed settling steps of the deposit")
    ap.add_argument("--max-iterations", type=int, default=100000,
                    help="APGD iteration cap of the timed steps (the reference's default, APGD.cu:11)")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only-mixed", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 512-scene batched solve")
    ap.add_argument("--no-serve", action="store_true", help="skip the served request stream")
    ap.add_argument("--no-shell", action="store_true", help="skip the spheres-on-a-fixed-shell leg")
    ap.add_argument("--only-shell", action="store_true",
                    help="run only the spheres-on-a-fixed-shell leg and print its figures")
    ap.add_argument("--only-fp64", action="store_true",
                    help="run only the FP64 graph-driven solver on the settled pile and print its figures")
    ap.add_argument("--only-batch", action="store_true",
                    help="run only the batched leg (with --batch-scenes scenes) and print its figures")
    ap.add_argument("--batch-scenes", type=int, default=512)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a 1 GiB copy kernel (for scaling rocprofv3 byte counters)")
    ap.add_argument("--summarize", metavar="CSV",
                    help="print a markdown summary of a rocprofv3 csv (kernel_stats or counter_collection) and exit")
    args = ap.parse_args()
    if args.summarize:
        summarize_rocprof(args.summarize)
        return
    if args.calibrate_copy:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the contact time step is not sharded across GPUs "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    from dvide_amd import scenes, system

    if args.only_batch:
        progress(f"{args.batch_scenes}-scene batch")
        print(json.dumps({"batched": batched(n_scenes=args.batch_scenes)}), flush=True)
        return
    if args.only_shell:
        progress("shell-bed leg")
        print(json.dumps({"shell_bed": shell_bed_leg()}), flush=True)
        return
    progress("settling the compaction pile")
    scene, settle_its = settled_scene(args.settle)
    if args.only_fp64:
        progress("FP64 graph-driven solver")
        print(json.dumps({"fp64": fp64_leg(scene, args.steps, args.warmup, args.max_iterations)}), flush=True)
        return
    progress("mixed solver")
    cap = args.max_iterations
    mixed = run_gpu(scene, system.BUILD_MIXED, args.steps, args.warmup, profile_kernel=system.PROFILE_P1,
                    max_iterations=cap)
    # burst-timed average of the body pass P2 (two untimed steps from the
    # same state); P1 + P2 is one solver pass
    kernel_us = {"P1": None, "P2": None}
    r = run_gpu(scene, system.BUILD_MIXED, 2, 0, profile_kernel=system.PROFILE_P2, max_iterations=cap)
    if r["kernel_count"]:
        kernel_us["P2"] = round(1e3 * r["kernel_ms"] / r["kernel_count"], 3)
    others = {}
    if not args.only_mixed:
        progress("FP64 graph-driven and plain FP64 solvers")
        # the schedule tuner measures each size class (log2 of the contact
        # count) once per device: an untimed FP64 step from the mixed leg's
        # final state, past 2^18 contacts like the FP64 leg's later steps,
        # keeps that measuring out of the timed steps
        late = copy.copy(scene)
        late.pos, late.vel = mixed["p"].reshape(-1, 3), mixed["v"].reshape(-1, 3)
        run_gpu(late, system.BUILD_FP64, 1, 0, max_iterations=cap)
        others["fp64"] = run_gpu(scene, system.BUILD_FP64, args.steps, args.warmup, max_iterations=cap)
        others["plain_fp64"] = run_gpu(scene, system.BUILD_PLAIN, max(2, args.steps // 4), args.warmup and 1,
                                       max_iterations=cap)

    value = args.steps / mixed["elapsed"]
    nc = int(np.mean(mixed["contacts"]))
    k_us = 1e3 * mixed["kernel_ms"] / max(mixed["kernel_count"], 1)
    # a pass after an accepted iterate forms N gamma_k (fresh), one after a
    # backtrack reads it: P1's bytes weighted by the share of fresh passes
    fresh = float(np.sum(mixed["iterations"])) / max(float(np.sum(mixed["passes"])), 1.0)
    b_p1 = fresh * p1_bytes_per_contact(True, True) + (1 - fresh) * p1_bytes_per_contact(True, False)
    alg = b_p1 * nc
    achieved = alg / (k_us * 1e-6) / 1e9 if k_us > 0 else 0.0
    # HBM bytes per P1 launch: the per-contact figure of the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE summary (scaled by a copy kernel) times this
    # workload's contacts
    traffic = None
    tpath = os.path.join(_HERE, "profiles", "apgd_traffic_compaction_112k_mixed.md")
    if os.path.exists(tpath):
        for ln in open(tpath):
            if ln.startswith("p1_traffic_bytes_per_contact:"):
                traffic = int(round(float(ln.split(":")[1]) * nc))

    def lat_ms(r):
        return dict(p50=round(1e3 * float(np.percentile(r["lat"], 50)), 3),
                    p99=round(1e3 * float(np.percentile(r["lat"], 99)), 3))

    line = {
        "metric": "time-steps/s at 250k contacts (res<=1e-4); p50/p99 solve latency, mixed vs FP64",
        "value": round(value, 3),
        "unit": "time-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * mixed["elapsed"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded scene generator)",
        "config": {
            "workload": "granular compaction: 112k spheres r~U[0.04,0.06] m in a 5x5 m box, settled "
                        f"{args.settle} steps, mu 0.3, h 1e-3, APGD tol 1e-4, warm start",
            "contacts": nc,
            "bodies": mixed["nb"],
            "solver": "APGD mixed (FP32 D/M^-1/products, FP64 iterates/tests/update)",
        },
        "latency_ms": lat_ms(mixed),
        "iterations_mean": round(float(np.mean(mixed["iterations"])), 1),
        "iterations_max": int(np.max(mixed["iterations"])),
        "residual_max": float(np.max(mixed["residual"])),
        "steps_converged": int(np.sum(mixed["residual"] < TOL)),
        "phases_ms_last_step": {k: round(v, 3) for k, v in mixed["phases"].items()},
        "schedule": mixed["schedule"],
        "roofline": {
            "bound": "hbm",
            "kernel": "kP1<float> (contact pass: previous decision, momentum, projection, F^T gamma)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "avg_us": round(k_us, 3),
            "launches_timed": mixed["kernel_count"],
            "avg_us_per_step": mixed["kernel_us_steps"],
            "bytes_per_contact": round(b_p1, 1),
            "fresh_pass_share": round(fresh, 4),
        },
    }
    kernel_us["P1"] = round(k_us, 3)
    line["kernel_us"] = kernel_us
    if kernel_us["P2"]:
        # one solver pass: P1 (the vector algebra, the stores of gamma_new and
        # F^T gamma_new, and the rows of N gamma_k) and P2 (the per-body
        # gather M^-1 D^T gamma_new and the Lipschitz term); the Schur product
        # is spread over the two
        b_pass = alg + p2_bytes(nc, mixed["nb"], True)
        t_pass = k_us + kernel_us["P2"]
        line["pass_phase"] = {
            "kernels": "kP1 + kBody (one APGD pass, the Schur product included)",
            "bytes": int(round(b_pass)),
            "us": round(t_pass, 3),
            "achieved_gbs": round(b_pass / (t_pass * 1e-6) / 1e9, 1),
            "frac": round(b_pass / (t_pass * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
        }
    def us_per_iteration(r):
        return round(1e3 * float(np.sum(r["solve_ms"])) / max(int(np.sum(r["iterations"])), 1), 2)

    line["us_per_iteration"] = us_per_iteration(mixed)
    for name, r in others.items():
        k = len(r["lat"])
        line[name] = {
            "value": round(k / r["elapsed"], 3),
            "ms_per_step": round(1e3 * r["elapsed"] / k, 3),
            "latency_ms": lat_ms(r),
            "iterations_mean": round(float(np.mean(r["iterations"])), 1),
            "residual_max": float(np.max(r["residual"])),
            "steps_converged": int(np.sum(r["residual"] < TOL)),
            "us_per_iteration": us_per_iteration(r),
            "schedule": r["schedule"],
        }
    # ratios over the same steps from the same state: the plain path runs the
    # first steps of the window only
    if "plain_fp64" in others:
        k = len(others["plain_fp64"]["lat"])
        line["mixed_vs_plain_fp64"] = {
            "time_per_step": round(float(np.sum(others["plain_fp64"]["lat"]) / np.sum(mixed["lat"][:k])), 2),
            "steps": k,
            "time_per_iteration": round(line["plain_fp64"]["us_per_iteration"] / line["us_per_iteration"], 2),
        }
    if "fp64" in others:
        line["mixed_vs_fp64"] = {
            "time_per_step": round(float(np.sum(others["fp64"]["lat"]) / np.sum(mixed["lat"])), 2),
            "steps": args.steps,
            "time_per_iteration": round(line["fp64"]["us_per_iteration"] / line["us_per_iteration"], 2),
        }
    line["settle_iterations"] = settle_its
    if not args.only_mixed:
        progress("lid-pressed leg")
        line["lid_pressed"] = lid_pressed()
    if not args.only_mixed and not args.no_shell:
        progress("shell-bed leg")
        line["shell_bed"] = shell_bed_leg()
    if not args.no_batch and not args.only_mixed:
        progress(f"{args.batch_scenes}-scene batch")
        line["batched"] = batched(n_scenes=args.batch_scenes)
    if not args.no_serve and not args.only_mixed:
        progress("served stream")
        line["served"] = served()
    if not args.no_cpu:
        # the CPU baseline leg: the oracle timed on the host cores (all of this
        # process's, then one), and, as the checker, the GPU solvers'
        # per-workload errors against it
        cores = host_cores()
        progress(f"CPU baseline on {cores} cores")
        base, first = cpu_baseline(scene, args.cpu_steps, cores)
        progress("CPU baseline on 1 core")
        one, _ = cpu_baseline(scene, 1, 1, warm_fill=False)
        base["cpu_model"] = cpu_model()
        base["single_thread"] = {"value": one["value"], "cores": 1, "sample": one["sample"]}
        line["cpu_baseline"] = base
        progress("error check against the oracle")
        oracle_threads(cores)
        line["error_vs_fp64_oracle"] = error_check(scene, first)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
