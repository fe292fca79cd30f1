#!/usr/bin/env python3
"""Throughput of the fused FP64 DG advection RK stage on one MI355X.

One "step" is one SSP-RK3 time step (three fused stage launches, replayed
from a hipGraph) of the periodic DG advection on the workload's mesh, with
the state resident in HBM.  The metric is GDOF-updates per second per RK
stage: Ne*Np*3*steps / seconds.

Prints one JSON line.  Beside the GPU figure it reports the stage kernel's
HBM roofline fraction (algorithmic bytes per DOF-update, averaged over the
three stages: 40 B at stage 1 -- read q,u,v,w, write q -- and 48 B at stages
2 and 3, which also read q0) and a CPU baseline: the reference's own Fortran
modules (oracle/_ref, compiled from source) timed on the host cores over a
bounded sample of the workload.

    python bench.py [--steps K] [--warmup W] [--workload c3_64cube_p7]
                    [--variant factored|exact|fma|...] [--no-cpu-baseline]
                    [--copy-calibration] [--no-error] [--no-extra]
                    [--ensemble-only | --side-only]

The line also carries the measured variant's rel-L2 and max error against
the FP64 oracle's bits after the workload's nstep, and (unless --no-extra)
the other workloads (8^3 p=3, 32^3 p=7, 16^3 p=15, the 64-instance ensemble)
and the components either side of the stage kernel on the workload's mesh
(device mesh build, diagnostics reductions, batched operator kernels).

--copy-calibration first copies 1 GiB device to device (hipMemcpy's copy
kernel) three times, so that a rocprofv3 --pmc run of this script also
records what FETCH_SIZE/WRITE_SIZE read for a copy of known size.

Unless --no-traffic or --no-extra, the HBM traffic per launch of every stage
kernel reported (roofline.traffic for the headline) is measured in the same
run: two child processes of this script under rocprofv3 --pmc (FETCH_SIZE,
then WRITE_SIZE) run one SSP-RK3 step of each variant with its tuning
variant, after three 1 GiB device copies that calibrate each counter.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "scale-dg-kernel-extraction_amd"))

import dgk  # noqa: E402
from dgk import ic  # noqa: E402

METRIC = "GDOF-updates/s per RK stage at p=7, 64^3 elems; rel-L2 err vs FP64 oracle"
HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per DOF-update, averaged over a step's three stages.
# Stage 1 reads q,u,v,w and writes q (40 B): the state it reads becomes q0
# without a copy.  Stages 2 and 3 also read q0 (48 B).
BYTES_PER_DOF = (40 + 48 + 48) / 3
# uniform velocity held as three scalars: 16 B at stage 1, 24 B after
BYTES_PER_DOF_CV = (16 + 24 + 24) / 3
# solver variants: flags of dgk_create_structured
VARIANTS = {
    "factored": dgk.FLAG_FACTORED_LIFT,                # FP64, lift by the 1D vectors
    "exact": 0,                                        # bit-identical to the reference
    "fma": dgk.FLAG_FMA,                               # fused multiply-adds
    "fma_factored": dgk.FLAG_FMA | dgk.FLAG_FACTORED_LIFT,
    "exact_node": dgk.FLAG_NODE_KERNEL,                # thread-per-node kernel
    # uniform velocity compressed to 3 scalars: 24 B per DOF-update
    "factored_constvel": dgk.FLAG_FACTORED_LIFT | dgk.FLAG_CONST_VELOCITY,
    "exact_constvel": dgk.FLAG_CONST_VELOCITY,
    "fma_node": dgk.FLAG_FMA | dgk.FLAG_NODE_KERNEL,
    "fma_factored_constvel": dgk.FLAG_FMA | dgk.FLAG_FACTORED_LIFT | dgk.FLAG_CONST_VELOCITY,
    # FP32 contractions of each line's remainder after its cubic part
    "mixed": dgk.FLAG_MIXED,
    "mixed_constvel": dgk.FLAG_MIXED | dgk.FLAG_CONST_VELOCITY,
    # p=15, uniform velocity: contractions on the matrix cores, FP64 MFMA, and
    # FP32 MFMA on each line's remainder after an FP64 degree-7 fit (the fit
    # and its derivative as FP64 MFMA products)
    "mfma_constvel": dgk.FLAG_MFMA | dgk.FLAG_CONST_VELOCITY,
    "mfma_mixed_constvel": dgk.FLAG_MFMA | dgk.FLAG_MIXED | dgk.FLAG_CONST_VELOCITY,
}
# MI355X dense matrix peaks (TFLOP/s)
MATRIX_PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def lgl_weights(x1D):
    p = len(x1D) - 1
    P = np.polynomial.legendre.Legendre.basis(p)(x1D)
    return 2.0 / (p * (p + 1) * P * P)


def mass(q, w3, J):
    """LGL-quadrature integral of q and of |q| over the domain."""
    return float(J * np.dot(w3, q.sum(axis=1))), float(J * np.dot(w3, np.abs(q).sum(axis=1)))


def host_cpu():
    """The host's CPU model, logical CPUs, sockets, NUMA nodes and the CPUs
    this process may run on."""
    model, sockets = None, set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model is None:
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("physical id"):
                    sockets.add(line.split(":", 1)[1].strip())
    except OSError:
        pass
    try:
        numa = len([d for d in os.listdir("/sys/devices/system/node") if d.startswith("node")
                    and d[4:].isdigit()])
    except OSError:
        numa = None
    return {"cpu_model": model, "logical_cpus": os.cpu_count(), "sockets": len(sockets) or None,
            "numa_nodes": numa, "cpus_allowed": len(os.sched_getaffinity(0))}


def cpu_baseline(threads, nstep, nrep, spec, q):
    """Times the reference's step loop (its own Fortran, compiled from source
    into oracle/_ref) with both of its operator types, GENERAL and OPT1:
    nrep runs of nstep steps of the workload itself (its mesh, order,
    velocity and dt) from the owned field q(Np,Ne), on `threads` OpenMP
    threads spread over cores (OMP_PROC_BIND=spread, OMP_PLACES=cores); the
    median per type.  Falls back to the C restatement (oracle/liboracle.so,
    OPT1 order) where the reference program is absent.  Only the step loop
    is timed (the reference's own loop), not its mesh setup or file
    exchange."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    ne, p, vel = spec["ne"], spec["p"], spec["vel"]
    u, v, w = (np.full_like(q, c) for c in vel)
    dt = ic.cfl_dt(p, vel, ne)
    dof_updates = q.size * 3 * nstep
    sample = (f"{ne[0]}x{ne[1]}x{ne[2]} elements, p={p}: {nrep} runs of {nstep} steps of the "
              f"workload per operator type ({dof_updates / 1e6:.1f} M DOF-updates a run), "
              f"median")
    env = {"OMP_PROC_BIND": "spread", "OMP_PLACES": "cores"}
    out = {"unit": "GDOF-updates/s", "cores": threads, "sample": sample, **host_cpu(),
           "omp": dict(env, OMP_NUM_THREADS=str(threads))}
    if O.ref_available():
        secs, used = O.ref_bench(*ne, p, q, u, v, w, dt, nstep, nrep, threads=threads, env=env)
        out["cores"] = used
        for t in ("GENERAL", "OPT1"):
            med = float(np.median(secs[t]))
            out[t] = {"value": dof_updates / med / 1e9, "seconds": secs[t]}
        out["kind"] = "reference"
    else:
        os.environ["OMP_NUM_THREADS"] = str(threads)
        om = O.Mesh(*ne, p)
        F = []
        for a in (q, u, v, w):
            f = om.field()
            f[:, :om.Ne] = a
            om.update_halo(f)
            F.append(f)
        secs = [om.run(*(x.copy(order="F") for x in F), dt, nstep, O.OPT1) for _ in range(nrep)]
        out["OPT1"] = {"value": dof_updates / float(np.median(secs)) / 1e9, "seconds": secs}
        out["kind"] = "port"
    # the headline baseline is the reference's default type, OPT1
    out["value"] = out["OPT1"]["value"]
    return out


STAGE_KERNELS = ("pencil_kernel", "pencil_group_kernel", "stage_kernel", "mfma16_kernel")


def traffic_probe(plan):
    """Child mode of measure_traffic, run under rocprofv3 --pmc: three 1 GiB
    device-to-device copies (the counters' calibration), then for each
    (workload, variant, tuning variant) of the plan a solver set up as in
    the timed runs and one SSP-RK3 step of stage launches (three dispatches
    of its stage kernel)."""
    import torch
    a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    fields = {}
    for workload, name, tv in plan:
        spec = ic.WORKLOADS[workload]
        ne, p, vel = spec["ne"], spec["p"], spec["vel"]
        x1D, D, L = dgk.read_operator(p)
        if workload not in fields:
            fields.clear()
            fields[workload] = _fields(spec, x1D)
        q = fields[workload]
        s = dgk.Solver.structured(*ne, p, D, L, flags=VARIANTS[name])
        s.set_fields(q, *(np.full_like(q, c) for c in vel))
        if tv is not None:
            s.set_variant(tv)
        s.synchronize()
        s.time_stages(ic.cfl_dt(p, vel, ne), 3)
        s.close()
        log(f"traffic probe: {workload} {name} (tuning variant {tv})")


def measure_traffic(plan, timeout=300):
    """HBM bytes per stage launch of each (workload, variant, tuning variant)
    in plan, measured in this run: two child processes of this script
    (traffic_probe) under rocprofv3 --pmc, one for FETCH_SIZE and one for
    WRITE_SIZE (KiB per dispatch), each scaled by what the counter reads for
    a 1 GiB copy in the same process.  Returns ({(workload, variant):
    record}, None) or (None, reason)."""
    import csv
    import shutil
    import signal
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None, "rocprofv3 not found"
    per = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="dgk_pmc_", dir="/tmp")
        cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "p", "--",
               sys.executable, os.path.abspath(__file__), "--traffic-probe", json.dumps(plan)]
        log(f"traffic: {counter} pass ({len(plan)} variants) under rocprofv3")
        proc = subprocess.Popen(cmd, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"),
                                stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                                start_new_session=True)
        try:
            rc = proc.wait(timeout=timeout)
        except subprocess.TimeoutExpired:
            os.killpg(proc.pid, signal.SIGKILL)
            proc.wait()
            shutil.rmtree(d, ignore_errors=True)
            return None, f"{counter} pass exceeded {timeout} s"
        rows = []
        for root, _, files in os.walk(d):
            for fn in files:
                if fn.endswith("counter_collection.csv"):
                    with open(os.path.join(root, fn)) as f:
                        rows += [r for r in csv.DictReader(f) if r["Counter_Name"] == counter]
        shutil.rmtree(d, ignore_errors=True)
        rows.sort(key=lambda r: int(r["Dispatch_Id"]))
        # the calibration copies are the probe's first copy dispatches (the
        # uploads that follow use the same copy kernel)
        copy = [float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"]][:3]
        stage = [float(r["Counter_Value"]) for r in rows
                 if any(k + "<" in r["Kernel_Name"] or k + "(" in r["Kernel_Name"]
                        for k in STAGE_KERNELS)]
        # rocprofv3 may end with a fault of its own after complete results
        if len(copy) < 3 or len(stage) != 3 * len(plan):
            return None, (f"{counter} pass: rc {rc}, {len(copy)} copy and {len(stage)} stage "
                          f"dispatches (expected 3 and {3 * len(plan)})")
        if max(copy) > 1.05 * min(copy) or min(copy) < (1 << 20) / 4:
            return None, f"{counter} pass: calibration copies read {copy} KiB"
        scale = (1 << 30) / (1024.0 * float(np.median(copy)))
        per[counter] = (scale, [1024.0 * v * scale for v in stage])
    out = {}
    (fs, fetch), (ws, write) = per["FETCH_SIZE"], per["WRITE_SIZE"]
    for i, (workload, name, tv) in enumerate(plan):
        f3, w3 = fetch[3 * i:3 * i + 3], write[3 * i:3 * i + 3]
        out[(workload, name)] = {
            "bytes_per_launch": round(sum(f3 + w3) / 3),
            "read_bytes_per_stage": [round(x) for x in f3],
            "write_bytes_per_stage": [round(x) for x in w3],
            "tuning_variant": tv, "fetch_scale": round(fs, 4), "write_scale": round(ws, 4)}
    return out, None


def _fields(spec, x1D):
    ne, p = spec["ne"], spec["p"]
    m = dgk.mesh_setup(*ne, p, x1D=x1D, tables=False)
    q = ic.initial_field(spec["ic"], m["pos_en"], spec["seed"])
    return q


def error_vs_oracle(workload, flags, x1D, D, L, spec):
    """rel-L2 and max error of the variant against the bit-exact variant
    after the workload's nstep, with the mass drift of both (device
    reductions)."""
    ne, p, vel, nstep = spec["ne"], spec["p"], spec["vel"], spec["nstep"]
    q = _fields(spec, x1D)
    dt = ic.cfl_dt(p, vel, ne)
    runs = []
    for f in (flags, VARIANTS["exact"]):
        s = dgk.Solver.structured(*ne, p, D, L, flags=f)
        s.set_fields(q, *(np.full_like(q, c) for c in vel))
        m0 = s.diagnostics(x1D)
        s.run(dt, nstep)
        runs.append((s, m0))
    (a, a0), (b, b0) = runs
    d = a.diagnostics(x1D, ref=b)
    db = b.diagnostics(x1D)
    out = {"nstep": nstep, "rel_l2": d["rel_l2"], "max_abs": d["max_err"],
           "max_rel": d["max_err"] / max(abs(db["min"]), abs(db["max"])),
           "mass_drift_rel": abs(d["mass"] - a0["mass"]) / a0["abs_mass"],
           "oracle_mass_drift_rel": abs(db["mass"] - b0["mass"]) / b0["abs_mass"],
           "reference": "exact variant (bit-identical to the FP64 oracle)"}
    a.close()
    b.close()
    return out


def mixed_workload(workload, x1D, D, L, spec, nstage=30):
    """DGK_FLAG_MIXED (FP32 contractions of the lines' remainders after
    their cubic part, FP64 elsewhere) on the workload: stage time with
    per-node and with uniform velocity, and its error against the
    bit-identical FP64 variant run alongside on the device, after every step
    of the workload's nstep (rel-L2, max error, mass drift)."""
    ne, p, vel = spec["ne"], spec["p"], spec["vel"]
    nstep = spec["nstep"]
    q = _fields(spec, x1D)
    dt = ic.cfl_dt(p, vel, ne)
    uvw = [np.full_like(q, c) for c in vel]
    res = {}
    for name, flags in (("per_node_velocity", VARIANTS["mixed"]),
                        ("uniform_velocity", VARIANTS["mixed_constvel"])):
        s = dgk.Solver.structured(*ne, p, D, L, flags=flags)
        s.set_fields(q, *uvw)
        var, _ = s.autotune(dt, nrep=6)
        ms = s.time_stages(dt, nstage)
        s.close()
        bpd = BYTES_PER_DOF_CV if flags & dgk.FLAG_CONST_VELOCITY else BYTES_PER_DOF
        dof = ne[0] * ne[1] * ne[2] * (p + 1) ** 3
        res[name] = {"gdof_updates_per_s": round(dof / (ms * 1e-3) / 1e9, 2),
                     "kernel_us": round(ms * 1e3, 1), "variant": var,
                     "hbm_frac": round(bpd * dof / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    a = dgk.Solver.structured(*ne, p, D, L, flags=VARIANTS["mixed"])
    b = dgk.Solver.structured(*ne, p, D, L, flags=VARIANTS["exact"])
    for x in (a, b):
        x.set_fields(q, *uvw)
    del q, uvw
    m0 = a.diagnostics(x1D)
    per_step = []
    for _ in range(nstep):
        a.step(dt)
        b.step(dt)
        a.synchronize()
        b.synchronize()
        per_step.append(float(f"{a.diagnostics(x1D, ref=b)['rel_l2']:.4e}"))
    d = a.diagnostics(x1D, ref=b)
    db = b.diagnostics(x1D)
    a.close()
    b.close()
    res["error_vs_fp64_oracle"] = {
        "nstep": nstep, "rel_l2": d["rel_l2"], "max_abs": d["max_err"],
        "max_rel": d["max_err"] / max(abs(db["min"]), abs(db["max"])),
        "mass_drift_rel": abs(d["mass"] - m0["mass"]) / m0["abs_mass"],
        "rel_l2_per_step": per_step,
        "reference": "exact variant (bit-identical to the FP64 oracle), stepped alongside"}
    res["variant"] = ("FP32 contractions (operator split hi+lo) of each line's remainder after "
                      "its cubic part, which is differentiated in FP64; lines whose remainder "
                      "exceeds 1/16 of the line in FP64; lift, flux differences, residual and "
                      "RK in FP64")
    log(f"mixed: {res['per_node_velocity']['gdof_updates_per_s']} GDOF-updates/s "
        f"(uniform velocity {res['uniform_velocity']['gdof_updates_per_s']}), rel-L2 "
        f"{d['rel_l2']:.2e} after {nstep} steps")
    return res


def flops_per_dof(p):
    """Algorithmic flops per DOF-update: three sweeps (2 Nq each), 22 for
    the flux products, combine, lift and RK, and the face fluxes (120/Nq)."""
    nq = p + 1
    return 6 * nq + 22 + 120 / nq


def single_workload(workload, variants, nstage=30, errors=False):
    """Stage kernel time of a few variants on another workload (after
    autotuning each), in GDOF-updates/s and TFLOP/s; with `errors`, each
    variant's rel-L2 against the bit-identical variant after the workload's
    nstep (both run on the device from the same field)."""
    spec = ic.WORKLOADS[workload]
    ne, p, vel = spec["ne"], spec["p"], spec["vel"]
    x1D, D, L = dgk.read_operator(p)
    q = _fields(spec, x1D)
    dt = ic.cfl_dt(p, vel, ne)
    uvw = [np.full_like(q, c) for c in vel]
    res = {"NeX": ne[0], "PolyOrder": p, "flops_per_dof_update": flops_per_dof(p)}
    ref = None
    if errors:
        # the bit-identical variant (the oracle's bits) after nstep: the
        # reference of every variant's rel-L2, max error and mass drift
        ref = dgk.Solver.structured(*ne, p, D, L, flags=VARIANTS["exact"])
        ref.set_fields(q, *uvw)
        r0 = ref.diagnostics(x1D)
        ref.run(dt, spec["nstep"])
        r1 = ref.diagnostics(x1D)
        ref_scale = max(abs(r1["min"]), abs(r1["max"]))
        res["oracle_mass_drift_rel"] = abs(r1["mass"] - r0["mass"]) / r0["abs_mass"]
    for name in variants:
        s = dgk.Solver.structured(*ne, p, D, L, flags=VARIANTS[name])
        s.set_fields(q, *uvw)
        var, _ = s.autotune(dt, nrep=6)
        ms = s.time_stages(dt, nstage)
        bpd = BYTES_PER_DOF_CV if VARIANTS[name] & dgk.FLAG_CONST_VELOCITY else BYTES_PER_DOF
        dofs = s.Ne * s.Np / (ms * 1e-3)
        r = {"gdof_updates_per_s": round(dofs / 1e9, 2), "kernel_us": round(ms * 1e3, 1),
             "variant": var, "hbm_frac": round(bpd * dofs / 1e9 / HBM_PEAK_GBS, 4),
             "tflops": round(flops_per_dof(p) * dofs / 1e12, 2),
             "fp64_peak_frac": round(flops_per_dof(p) * dofs / 1e12 / MATRIX_PEAK_TFLOPS["fp64"], 4)}
        if VARIANTS[name] & dgk.FLAG_MFMA:
            # flops the matrix cores execute per DOF-update: the three sweeps,
            # 2 Nq each (FP64; in the mixed form FP32 on the remainders), and
            # in the mixed form the FP64 fit and correction products, 2 x 8
            # nodes x 2 per direction; the fraction is of the matrix cores'
            # time at their dense peaks (78.6 FP64, 157.3 FP32 TFLOP/s)
            mixed = bool(VARIANTS[name] & dgk.FLAG_MIXED)
            f32 = 6 * (p + 1) if mixed else 0
            f64 = 3 * 2 * 8 * 2 if mixed else 6 * (p + 1)
            r.update(matrix_tflops_executed_fp64=round(f64 * dofs / 1e12, 2),
                     matrix_tflops_executed_fp32=round(f32 * dofs / 1e12, 2),
                     matrix_peak_tflops=MATRIX_PEAK_TFLOPS,
                     matrix_peak_frac=round(f64 * dofs / 1e12 / MATRIX_PEAK_TFLOPS["fp64"] +
                                            f32 * dofs / 1e12 / MATRIX_PEAK_TFLOPS["fp32"], 4))
        if ref is not None:
            s.set_fields(q, *uvw)
            m0 = s.diagnostics(x1D)
            s.run(dt, spec["nstep"])
            d = s.diagnostics(x1D, ref=ref)
            r["rel_l2_vs_exact_after_nstep"] = d["rel_l2"]
            r["error_vs_fp64_oracle"] = {
                "nstep": spec["nstep"], "rel_l2": d["rel_l2"], "max_abs": d["max_err"],
                "max_rel": d["max_err"] / ref_scale,
                "mass_drift_rel": abs(d["mass"] - m0["mass"]) / m0["abs_mass"]}
        res[name] = r
        s.close()
        log(f"{workload} {name}: {r['gdof_updates_per_s']} GDOF-updates/s")
    if ref is not None:
        ref.close()
        res["nstep"] = spec["nstep"]
    return res


def ensemble_workload(nstage=30):
    """The 64-instance ensemble (orders 3, 5, 7, 11, uneven meshes, uniform
    per-instance velocity): one persistent launch per stage, tile shape and
    grid split tuned per order class (the untuned default shape beside it); FP64 with fused multiply-adds, and the mixed
    form with its error against the FP64 ensemble after the workload's 50
    steps."""
    inst = ic.ensemble_workload(**{k: v for k, v in ic.ENSEMBLE.items() if k in ("n", "seed")})
    fields = []
    for sp in inst:
        x1D, _, _ = dgk.read_operator(sp["p"])
        m = dgk.mesh_setup(*sp["ne"], sp["p"], x1D=x1D, tables=False)
        fields.append(ic.blob(m["pos_en"], sp["centre"], sp["width"]))

    def build(flags):
        ens = dgk.Ensemble(inst, flags=flags)
        for i, sp in enumerate(inst):
            qi = fields[i]
            ens.set_fields(i, qi, *(np.full_like(qi, c) for c in sp["vel"]))
        ens.set_dt([sp["dt"] for sp in inst])
        ens.tune(3)
        return ens

    ens = build(dgk.FLAG_CONST_VELOCITY | dgk.FLAG_FACTORED_LIFT | dgk.FLAG_FMA)
    ms = ens.time_stages(nstage)
    t0 = time.perf_counter()
    ens.run(5)
    el = time.perf_counter() - t0
    dof = ens.query(0)
    nt, epb = ens.shape
    res = {"instances": len(inst), "DOF": dof, "classes": ens.classes,
           "tile_shape": {"threads": nt, "elems_per_item": epb},
           "gdof_updates_per_s_kernel": round(dof / (ms * 1e-3) / 1e9, 2),
           "gdof_updates_per_s_graph": round(dof * 15 / el / 1e9, 2),
           "variant": "FP64, fused multiply-adds, factored lift, uniform velocity per instance"}
    ens.set_shape()  # the default shape, grid split by DOF
    ms0 = ens.time_stages(nstage)
    res["untuned"] = {"gdof_updates_per_s_kernel": round(dof / (ms0 * 1e-3) / 1e9, 2),
                      "tile_shape": {"threads": ens.shape[0], "elems_per_item": ens.shape[1]},
                      "split": "by DOF"}
    ens.close()
    ens = build(dgk.FLAG_CONST_VELOCITY | dgk.FLAG_MIXED)
    ms = ens.time_stages(nstage)
    res["mixed"] = {"gdof_updates_per_s_kernel": round(dof / (ms * 1e-3) / 1e9, 2),
                    "tile_shape": {"threads": ens.shape[0], "elems_per_item": ens.shape[1]},
                    "variant": "FP32 contractions of the lines' remainders after their cubic "
                               "part, FP64 elsewhere, uniform velocity"}
    ens.close()
    # the measured form and the mixed form against the bit-identical FP64
    # ensemble after the workload's nstep, per instance: rel-L2, max error and
    # the drift of the LGL-quadrature mass of each instance
    nstep = ic.ENSEMBLE["nstep"]
    b = build(dgk.FLAG_CONST_VELOCITY)
    b.run(nstep)
    ref = [b.get_field(i) for i in range(len(inst))]
    b.close()
    quad = []
    for sp in inst:
        w3 = np.einsum("i,j,k->ijk", *(lgl_weights(dgk.read_operator(sp["p"])[0]),) * 3)
        quad.append((w3.ravel(order="F"),
                     (1.0 / sp["ne"][0]) * (1.0 / sp["ne"][1]) * (1.0 / sp["ne"][2]) / 8.0))

    def errors_of(flags):
        a = build(flags)
        a.run(nstep)
        errs, num, den, mx, mx_rel, drift, drift_ref = [], 0.0, 0.0, 0.0, 0.0, 0.0, 0.0
        for i in range(len(inst)):
            x, y = ref[i], a.get_field(i)
            d2, r2 = float(np.sum((y - x) ** 2)), float(np.sum(x * x))
            errs.append((d2 / r2) ** 0.5)
            num += d2
            den += r2
            e = float(np.max(np.abs(y - x)))
            mx = max(mx, e)
            mx_rel = max(mx_rel, e / float(np.max(np.abs(x))))
            w3, J = quad[i]
            m0, a0 = mass(fields[i], w3, J)
            drift = max(drift, abs(mass(y, w3, J)[0] - m0) / a0)
            drift_ref = max(drift_ref, abs(mass(x, w3, J)[0] - m0) / a0)
        a.close()
        return {"nstep": nstep, "max_instance_rel_l2": max(errs),
                "median_instance_rel_l2": float(np.median(errs)),
                "ensemble_rel_l2": (num / den) ** 0.5, "max_abs": mx,
                "max_instance_max_rel": mx_rel, "max_instance_mass_drift_rel": drift,
                "oracle_max_instance_mass_drift_rel": drift_ref,
                "reference": "the exact FP64 ensemble (bit-identical per instance to the oracle)"}

    res["error_vs_fp64_oracle"] = errors_of(
        dgk.FLAG_CONST_VELOCITY | dgk.FLAG_FACTORED_LIFT | dgk.FLAG_FMA)
    res["mixed"]["error_vs_fp64"] = errors_of(dgk.FLAG_CONST_VELOCITY | dgk.FLAG_MIXED)
    log(f"ensemble: {res['gdof_updates_per_s_kernel']} GDOF-updates/s per stage "
        f"(mixed {res['mixed']['gdof_updates_per_s_kernel']})")
    return res


def host_inclusive(spec, flags, q, nstep):
    """The whole job as a host caller sees it through the C-ABI: upload of
    q, u, v, w from pageable host arrays (dgk_set_fields), nstep steps
    (dgk_run) and the download of q (dgk_get_field), by the host clock."""
    ne, p, vel = spec["ne"], spec["p"], spec["vel"]
    _, D, L = dgk.read_operator(p)
    dt = ic.cfl_dt(p, vel, ne)
    s = dgk.Solver.structured(*ne, p, D, L, flags=flags)
    # the host arrays a caller of the C-ABI holds: (Np,NeA), the halo
    # columns included (the Python binding would pad (Np,Ne) arrays itself,
    # a host copy that is not part of the transfer)
    q = s._full(q)
    uvw = [np.full_like(q, c) for c in vel]
    s.set_fields(q, *uvw)  # first touch of the device buffers, not timed
    s.run(dt, 1)
    t0 = time.perf_counter()
    s.set_fields(q, *uvw)
    t1 = time.perf_counter()
    s.run(dt, nstep)
    t2 = time.perf_counter()
    s.get_field()
    t3 = time.perf_counter()
    dof = s.Np * s.Ne
    s.close()
    return {"nstep": nstep, "upload_s": round(t1 - t0, 3), "steps_s": round(t2 - t1, 3),
            "download_s": round(t3 - t2, 3),
            "gdof_updates_per_s": round(dof * 3 * nstep / (t3 - t0) / 1e9, 2),
            "bytes": "q, u, v, w up from (Np,NeA) pageable host arrays, staged through "
                     "pinned buffers by several host threads, and q down the same way"}


def _hbm_row(us, nbytes, what):
    gbs = nbytes / (us * 1e-6) / 1e9
    return {"us": round(us, 1), "algorithmic_bytes": int(nbytes), "bytes": what,
            "GBps": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}


class _Hip:
    """The HIP runtime that libdgk.so is linked against, opened by its soname
    (so the process keeps that one runtime): device buffers, a stream and
    events for the side measurements."""

    def __init__(self):
        import ctypes as C
        self.C = C
        dgk.lib()
        self.h = C.CDLL("libamdhip64.so.7")
        vp, pvp = C.c_void_p, C.POINTER(C.c_void_p)
        for name, args in (("hipMalloc", [pvp, C.c_size_t]), ("hipFree", [vp]),
                           ("hipMemset", [vp, C.c_int, C.c_size_t]),
                           ("hipMemcpy", [vp, vp, C.c_size_t, C.c_int]),
                           ("hipStreamCreate", [pvp]), ("hipStreamDestroy", [vp]),
                           ("hipStreamSynchronize", [vp]),
                           ("hipEventCreate", [pvp]), ("hipEventDestroy", [vp]),
                           ("hipEventRecord", [vp, vp]), ("hipEventSynchronize", [vp]),
                           ("hipEventElapsedTime", [C.POINTER(C.c_float), vp, vp])):
            f = getattr(self.h, name)
            f.argtypes, f.restype = args, C.c_int
        self.bufs = []
        st = vp()
        self.ok(self.h.hipStreamCreate(C.byref(st)), "hipStreamCreate")
        self.stream = st.value

    def ok(self, rc, what):
        if rc:
            raise RuntimeError(f"{what} returned HIP error {rc}")

    def zeros(self, nbytes):
        p = self.C.c_void_p()
        self.ok(self.h.hipMalloc(self.C.byref(p), nbytes), "hipMalloc")
        self.bufs.append(p.value)
        self.ok(self.h.hipMemset(p.value, 0, nbytes), "hipMemset")
        return p.value

    def upload(self, a):
        a = np.asfortranarray(a, dtype=np.float64)
        p = self.zeros(a.nbytes)
        self.ok(self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1), "hipMemcpy")
        return p

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []

    def close(self):
        self.free()
        self.h.hipStreamDestroy(self.stream)

    def time_us(self, fn, nrep):
        """Mean time of fn() (which launches on self.stream) between two
        events recorded on that stream."""
        C = self.C
        fn()
        self.ok(self.h.hipStreamSynchronize(self.stream), "hipStreamSynchronize")
        e0, e1 = C.c_void_p(), C.c_void_p()
        self.ok(self.h.hipEventCreate(C.byref(e0)), "hipEventCreate")
        self.ok(self.h.hipEventCreate(C.byref(e1)), "hipEventCreate")
        self.ok(self.h.hipEventRecord(e0, self.stream), "hipEventRecord")
        for _ in range(nrep):
            fn()
        self.ok(self.h.hipEventRecord(e1, self.stream), "hipEventRecord")
        self.ok(self.h.hipEventSynchronize(e1), "hipEventSynchronize")
        ms = C.c_float(0.0)
        self.ok(self.h.hipEventElapsedTime(C.byref(ms), e0, e1), "hipEventElapsedTime")
        self.h.hipEventDestroy(e0)
        self.h.hipEventDestroy(e1)
        return ms.value / nrep * 1e3


def streaming_reference(gib=4, nrep=20):
    """What a plain streaming kernel moves on this GPU, for scale beside the
    stage kernel: torch.addcmul (three reads, one write of `gib` GiB each)
    and a copy, timed with events on torch's stream."""
    import torch
    n = (gib << 30) // 8
    a, b, c = (torch.ones(n, dtype=torch.float64, device="cuda") for _ in range(3))
    o = torch.empty_like(a)
    out = {}
    for name, fn, nbytes in (("addcmul_3r1w", lambda: torch.addcmul(c, a, b, out=o), 4),
                             ("copy_1r1w", lambda: o.copy_(a), 2)):
        for _ in range(3):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(nrep):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / nrep
        gbs = nbytes * (gib << 30) / (ms * 1e-3) / 1e9
        out[name] = {"GBps": round(gbs, 1), "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4)}
    del a, b, c, o
    torch.cuda.empty_cache()
    out["arrays"] = f"{gib} GiB FP64 each, torch kernels, HIP events over {nrep} launches"
    return out


def side_rows(spec, q, nrep=5):
    """The components either side of the stage kernel, on the workload's
    mesh: the device mesh build (against the host build of the same tables),
    the diagnostics reductions and the batched OPT1 operator kernels over
    every element (inputs zero-filled: the kernels' time does not depend on
    the values).  Each is HBM-bound; the row gives its time per call (HIP
    events on the stream it is launched on, or the host clock around a call
    that returns to the host), its algorithmic bytes and their fraction of
    the 8 TB/s peak."""
    ne, p, vel = spec["ne"], spec["p"], spec["vel"]
    x1D, D, L = dgk.read_operator(p)
    sz = dgk.mesh_sizes(*ne, p)
    Nq, Np, Nf, Ne, nhn = sz["Nq"], sz["Np"], sz["NfpTot"], sz["Ne"], sz["NhaloNode"]
    hip = _Hip()
    sh = hip.stream

    def clock(fn):
        fn()
        t0 = time.perf_counter()
        for _ in range(nrep):
            fn()
        return (time.perf_counter() - t0) / nrep * 1e6

    out = {"NeX": ne[0], "PolyOrder": p}
    # device mesh build: every table of mesh_setup, written once
    sizes = {"VMapM": 4 * Nf * Ne, "VMapP": 4 * Nf * Ne, "normal_fn": 24 * Nf * Ne,
             "Escale": 24 * Np * Ne, "Fscale": 8 * Nf * Ne, "pos_en": 24 * Np * Ne,
             "halo_src_map": 4 * nhn}
    ptrs = {k: hip.zeros(n) for k, n in sizes.items()}
    row = _hbm_row(hip.time_us(lambda: dgk.mesh_setup_device(*ne, p, x1D, ptrs, stream=sh),
                               nrep),
                   sum(sizes.values()), "tables written: VMapM, VMapP, normal_fn, Escale, "
                                        "Fscale, pos_en, halo_src_map")
    hip.free()
    t0 = time.perf_counter()
    m = dgk.mesh_setup(*ne, p, x1D=x1D)
    row["host_build_s"] = round(time.perf_counter() - t0, 2)
    del m
    out["mesh_setup_device"] = row
    log(f"mesh build: device {row['us'] / 1e3:.1f} ms, host {row['host_build_s']} s")

    # diagnostics: min/max/mass/sum q^2, and rel-L2 against a second field
    s = dgk.Solver.structured(*ne, p, D, L, flags=VARIANTS["factored"])
    s.set_fields(q, *(np.full_like(q, c) for c in vel))
    ref = s.field_ptr(dgk.FIELD_U)
    out["diagnostics"] = _hbm_row(clock(lambda: s.diagnostics(x1D)), 8 * Np * Ne,
                                  "q read once (8 B per DOF); host clock around the call")
    out["diagnostics_vs_ref"] = _hbm_row(clock(lambda: s.diagnostics(x1D, ref=ref)),
                                         16 * Np * Ne, "q and the reference field read once")
    dt = ic.cfl_dt(p, vel, ne)
    nstep = 10
    s.run(dt, 2)
    t0 = time.perf_counter()
    s.run(dt, nstep)
    plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    s.run(dt, nstep, minmax=True)
    with_mm = time.perf_counter() - t0
    out["step_minmax"] = {"us_per_step_added": round((with_mm - plain) / nstep * 1e6, 1),
                          "bytes": "q read once per step (8 B per DOF)",
                          "algorithmic_bytes": 8 * Np * Ne}
    s.close()
    log(f"diagnostics: {out['diagnostics']['us']:.0f} us, with ref "
        f"{out['diagnostics_vs_ref']['us']:.0f} us")

    # batched OPT1 operator kernels over every element of the mesh
    n = Ne
    dD, dDt = hip.upload(D), hip.upload(D.T)
    dL = hip.upload(L)
    ip = [hip.zeros(8 * Np * n) for _ in range(3)]
    op = [hip.zeros(8 * Np * n) for _ in range(3)]
    fb = hip.zeros(8 * Nf * n)
    out["divlike_dirXYZ"] = _hbm_row(
        hip.time_us(lambda: dgk.divlike_batched(Nq, n, dgk.TYPE_OPT1, dD, dDt, *ip, *op,
                                                stream=sh), nrep),
        48 * Np * n, "3 fields in, 3 out")
    out["gradlike_dirXYZ"] = _hbm_row(
        hip.time_us(lambda: dgk.gradlike_batched(Nq, n, dD, dDt, ip[0], *op, stream=sh), nrep),
        32 * Np * n, "1 field in, 3 out")
    out["matvec_dirX"] = _hbm_row(
        hip.time_us(lambda: dgk.matvec_dir_batched(Nq, n, 0, dD, ip[0], op[0], stream=sh),
                    nrep),
        16 * Np * n, "1 field in, 1 out")
    out["lift_hexahedral"] = _hbm_row(
        hip.time_us(lambda: dgk.lift_batched(Nq, n, dL, fb, op[0], stream=sh), nrep),
        8 * (Nf + Np) * n, "face fluxes in (NfpTot per element), 1 field out")
    hip.close()
    log("operators: " + ", ".join(f"{k} {out[k]['us']:.0f} us" for k in
                                  ("divlike_dirXYZ", "gradlike_dirXYZ", "matvec_dirX",
                                   "lift_hexahedral")))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c3_64cube_p7", choices=sorted(ic.WORKLOADS))
    ap.add_argument("--variant", default="factored", choices=sorted(VARIANTS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 passes that measure HBM traffic per launch")
    ap.add_argument("--traffic-probe", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--no-error", action="store_true",
                    help="skip the nstep error run against the bit-exact variant")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the p=15 and ensemble side measurements")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="measure only the 64-instance ensemble (one JSON line)")
    ap.add_argument("--side-only", action="store_true",
                    help="measure only the rows beside the stage kernel on the workload's "
                         "mesh (one JSON line)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "sharding of the mesh is a design, not a measured path", file=sys.stderr)
        sys.exit(2)

    # torch's wheel carries its own HIP and HSA runtimes beside the ones
    # libdgk.so links: torch's must take the GPU first (as in the tests), or
    # its later initialisation finds no device once libdgk's has it
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU (torch.cuda.is_available() is False)", file=sys.stderr)
        sys.exit(2)
    torch.cuda.init()

    if args.traffic_probe:
        traffic_probe([tuple(x) for x in json.loads(args.traffic_probe)])
        return
    if args.ensemble_only:
        print(json.dumps({"workload": "c4_ensemble_64", **ensemble_workload()}), flush=True)
        return
    if args.side_only:
        spec = ic.WORKLOADS[args.workload]
        q = _fields(spec, dgk.read_operator(spec["p"])[0])
        print(json.dumps({"workload": args.workload, **side_rows(spec, q)}), flush=True)
        return

    if args.copy_calibration:
        import torch
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
        log("copied 1 GiB device-to-device 3 times (counter calibration)")

    spec = ic.WORKLOADS[args.workload]
    ne, p, vel = spec["ne"], spec["p"], spec["vel"]
    log(f"workload {args.workload}: {ne} elements, p={p}, variant {args.variant}")
    x1D, D, L = dgk.read_operator(p)
    t0 = time.perf_counter()
    m = dgk.mesh_setup(*ne, p, x1D=x1D, tables=False)
    q = ic.initial_field(spec["ic"], m["pos_en"], spec["seed"])
    del m
    u, v, w = (np.full_like(q, c) for c in vel)
    dt = ic.cfl_dt(p, vel, ne)
    log(f"initial field generated in {time.perf_counter() - t0:.1f} s")
    s = dgk.Solver.structured(*ne, p, D, L, flags=VARIANTS[args.variant])
    if s.kernel_kind != "pencil":
        kernel_name = "stage_kernel"
    elif p + 1 not in (2, 4, 8) or s.variant & 4:
        kernel_name = "pencil_group_kernel"
    else:
        kernel_name = "pencil_kernel"
    bytes_per_dof = (BYTES_PER_DOF_CV if VARIANTS[args.variant] & dgk.FLAG_CONST_VELOCITY
                     else BYTES_PER_DOF)
    s.set_fields(q, u, v, w)
    w3 = np.einsum("i,j,k->ijk", *(lgl_weights(x1D),) * 3).ravel(order="F")
    J = (1.0 / ne[0]) * (1.0 / ne[1]) * (1.0 / ne[2]) / 8.0
    mass0, mass_abs = mass(q, w3, J)
    del q, u, v, w
    Ne, Np = s.Ne, s.Np

    # the reference program's step loop (dgk_run: three stage launches per
    # step on the solver's stream; it returns once the steps are done)
    s.run(dt, args.warmup)
    s.synchronize()
    log(f"warmup done ({args.warmup} steps)")
    t0 = time.perf_counter()
    s.run(dt, args.steps)
    s.synchronize()
    el = time.perf_counter() - t0
    dof = Ne * Np
    value = dof * 3 * args.steps / el / 1e9
    log(f"timed {args.steps} steps in {el:.4f} s -> {value:.2f} GDOF-updates/s")

    # dominant kernel: the fused stage launch, timed with HIP events on the
    # solver's own stream
    nstage = max(30, 3 * args.steps)
    ms_stage = s.time_stages(dt, nstage)
    bytes_launch = round(bytes_per_dof * dof)  # mean over the three stages
    achieved = bytes_launch / (ms_stage * 1e-3) / 1e9
    qf = s.get_field()
    finite = bool(np.isfinite(qf).all())
    # the bit-identical variant's stage kernel, timed on the same state
    ms_exact = None
    tuning = {args.variant: s.variant}
    if args.variant != "exact":
        se = dgk.Solver.structured(*ne, p, D, L, flags=VARIANTS["exact"])
        se.set_fields(qf, *(np.full_like(qf, c) for c in vel))
        se.run(dt, 1)
        ms_exact = se.time_stages(dt, nstage)
        tuning["exact"] = se.variant
        se.close()
    # the sine-mode field has (near) zero mean: drift is measured against the
    # integral of |q|
    drift = abs(mass(qf, w3, J)[0] - mass0) / mass_abs
    s.close()
    log(f"stage kernel {ms_stage * 1e3:.1f} us/launch -> {achieved:.0f} GB/s algorithmic")

    # error against the FP64 oracle after the workload's nstep: the exact
    # variant gives the oracle's bits (the parity tests show it), so the
    # measured variant is compared with it on the device, from the same
    # initial field
    err = None
    if not args.no_error:
        err = error_vs_oracle(args.workload, VARIANTS[args.variant], x1D, D, L, spec)
        log(f"after {err['nstep']} steps: rel-L2 {err['rel_l2']:.3e}, max {err['max_abs']:.3e}")

    extra = {}
    if not args.no_extra:
        # the same job with its host transfers (the C-ABI hands over host arrays)
        extra["host_inclusive_" + args.workload] = host_inclusive(spec, VARIANTS[args.variant],
                                                                  qf, spec["nstep"])
        log("host-inclusive: " + json.dumps(extra["host_inclusive_" + args.workload]))
        extra["c3_64cube_p7_mixed"] = mixed_workload(args.workload, x1D, D, L, spec)
        # the same mesh with the uniform velocity held as three scalars, and
        # the FMA form with per-node velocity (each autotuned)
        extra["c3_64cube_p7_other_variants"] = single_workload(
            "c3_64cube_p7", ("factored_constvel", "fma_factored_constvel", "fma_factored"))
        # the high-order configuration: FP64 on the vector units (exact,
        # factored, FMA) against FP64 MFMA and FP32-split MFMA contractions
        extra["c5_16cube_p15"] = single_workload(
            "c5_16cube_p15", ("exact", "factored", "factored_constvel", "fma_factored_constvel",
                              "mfma_constvel", "mfma_mixed_constvel", "mixed"), errors=True)
        # the two smaller configurations: the reference's CPU-sized case and
        # the 32^3 mesh of the headline order
        extra["c1_8cube_p3"] = single_workload("c1_8cube_p3", ("exact", "factored"))
        extra["c2_32cube_p7"] = single_workload("c2_32cube_p7", ("exact", "factored"))
        extra["c4_ensemble_64"] = ensemble_workload()
        # the rows either side of the stage kernel, on the same mesh
        extra["side_rows_" + args.workload] = side_rows(spec, qf)

    # HBM traffic per launch of every stage kernel reported above, measured
    # in this run (rocprofv3 passes in child processes)
    traffic, traffic_note = None, "not measured (--no-traffic or --no-extra)"
    if not (args.no_traffic or args.no_extra):
        rows = [(args.workload, name, tv, None) for name, tv in tuning.items()]
        mx = extra["c3_64cube_p7_mixed"]
        rows += [("c3_64cube_p7", "mixed", mx["per_node_velocity"]["variant"],
                  mx["per_node_velocity"]),
                 ("c3_64cube_p7", "mixed_constvel", mx["uniform_velocity"]["variant"],
                  mx["uniform_velocity"])]
        for key, workload in (("c3_64cube_p7_other_variants", "c3_64cube_p7"),
                              ("c5_16cube_p15", "c5_16cube_p15"), ("c2_32cube_p7", "c2_32cube_p7"),
                              ("c1_8cube_p3", "c1_8cube_p3")):
            for name in VARIANTS:
                if isinstance(extra[key].get(name), dict):
                    rows.append((workload, name, extra[key][name]["variant"], extra[key][name]))
        traffic, traffic_note = measure_traffic([r[:3] for r in rows])
        if traffic is None:
            log(f"traffic not measured: {traffic_note}")
        else:
            traffic_note = ("rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE in child processes of this "
                            "run, one SSP-RK3 step per variant, each counter scaled by its "
                            "reading for a 1 GiB device copy")
            for workload, name, _, rec in rows:
                tr = traffic[(workload, name)]
                sp = ic.WORKLOADS[workload]
                tr["algorithmic_bytes_per_launch"] = round(
                    (BYTES_PER_DOF_CV if VARIANTS[name] & dgk.FLAG_CONST_VELOCITY
                     else BYTES_PER_DOF) * sp["ne"][0] * sp["ne"][1] * sp["ne"][2] *
                    (sp["p"] + 1) ** 3)
                tr["vs_algorithmic"] = round(tr["bytes_per_launch"] /
                                             tr["algorithmic_bytes_per_launch"], 4)
                if rec is not None:
                    rec["traffic_bytes_per_launch"] = tr["bytes_per_launch"]
                    rec["traffic_vs_algorithmic"] = tr["vs_algorithmic"]
            log("traffic measured for " + ", ".join(f"{w}/{n}" for w, n, _, _ in rows))
    head_traffic = traffic.get((args.workload, args.variant)) if traffic else None
    exact_traffic = traffic.get((args.workload, "exact")) if traffic else None

    stream_ref = None if args.no_extra else streaming_reference()

    cpu = None
    if not args.no_cpu_baseline:
        # the CPUs this job may use: the OpenMP share the host gives it (16
        # per GPU on the GPU boxes, where OMP_NUM_THREADS is set), else every
        # CPU it is allowed to run on
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS") or 0) or \
            len(os.sched_getaffinity(0))
        log(f"cpu baseline on {threads} threads")
        cpu = cpu_baseline(threads, args.cpu_steps, args.cpu_reps, spec, qf)
        log(f"cpu baseline OPT1 {cpu['value']:.4f} GDOF-updates/s ({cpu['kind']})")

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "GDOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded sine-mode initial field, constant velocity stored per node)",
        "config": {"workload": args.workload, "NeX": ne[0], "NeY": ne[1], "NeZ": ne[2],
                   "PolyOrder": p, "velocity": list(vel), "dt": dt, "variant": args.variant,
                   "DOF": dof, "dof_updates_per_step": 3 * dof},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": head_traffic["bytes_per_launch"] if head_traffic else None,
                     "kernel": kernel_name, "kernel_us": round(ms_stage * 1e3, 2),
                     "algorithmic_bytes_per_launch": bytes_launch,
                     "traffic_vs_algorithmic": head_traffic["vs_algorithmic"] if head_traffic
                     else None,
                     "traffic_detail": head_traffic, "traffic_source": traffic_note,
                     "streaming_reference": stream_ref},
        "cpu_baseline": cpu,
        "checks": {"finite": finite, "mass_drift_rel": drift},
        "error_vs_fp64_oracle": err,
        "other_workloads": extra or None,
        "exact_variant": None if ms_exact is None else {
            "kernel_us": round(ms_exact * 1e3, 2),
            "gdof_updates_per_s": round(dof / (ms_exact * 1e-3) / 1e9, 2),
            # the bit-identical variant streams per-node velocity
            "hbm_frac": round(BYTES_PER_DOF * dof / (ms_exact * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "traffic_bytes_per_launch": exact_traffic["bytes_per_launch"] if exact_traffic
            else None,
            "traffic_vs_algorithmic": exact_traffic["vs_algorithmic"] if exact_traffic else None},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
