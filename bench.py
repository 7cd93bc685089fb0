"""Throughput of the distance-field path on one MI355X.

Default workload (the production shape the metric is quoted on): bumpy icosphere split 8 times
(1,310,720 triangles, 32 bumps); one step = a 256^3 signed-distance volume (calcFromGridWithDistances,
FWN beta 2) + findProjections of 2^22 query points (half uniform in the bbox grown 20 %, half in a
Gaussian shell of sigma = 2 % of the bbox diagonal).  Inputs are resident in HBM before the timed
region; HIP events on the launch stream time every call.

Prints ONE JSON line: value = (voxels + projected points) per second for the whole step, with
* the per-part rates and one step of the per-lane (reference-form) kernels, outputs identical;
* `host_pointer`: the same two calls through the host-array entry points (PCIe copies included);
* the learned bound (an MLP trained on the GPU, capped search at three margins, outputs identical)
  against the plain walk, and the ceiling any bound can reach (caps = the true distances);
* `batched`: the uneven 64-mesh workload in one persistent launch (plain and learned);
* `offset`: the 512^3 volume of a 5,242,880-triangle mesh (plain, learned, and the band options of
  the offset caller on the device path and through calcFromGridWithDistancesByParts);
* `roofline` (the dominant call: the signed-distance grid) and `rooflines` (every call): achieved =
  the bytes of the records the walks read, counted per wave by the kernels themselves (each record a
  wave reads once for its 64 lanes counts once), divided by the call's event time, against the HBM
  peak; MLP kernels against the BF16 MFMA peak.  `traffic` / L2 hit rate come from the committed PMC
  profile (profiles/pmc_*.json) when it was taken on the same kernel sources, else null;
* `cpu_baseline`: the oracle (a C restatement of the reference CPU path) on the host cores available
  to this process, over contiguous z-slabs / query ranges of the same workload, median of 5 after a
  warm-up.

Other modes: --profile-run (one call of each kernel family, for rocprofv3), --summarize-pmc DIR NAME
(rocprofv3 counter CSVs -> profiles/pmc_NAME.json).
"""
import argparse
import csv
import glob
import hashlib
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
MESH_LEVEL, MESH_BUMPS, MESH_SEED = 8, 32, 2
GRID = 256
NQ = 1 << 22
QSEED = 7
TRAIN_N, TRAIN_BATCH, TRAIN_LR, TRAIN_SEED = 1 << 20, 8192, 1e-3, 11
MARGINS = (0.05, 0.1, 0.25)
CAL_N, CAL_SEED, CAL_Q = 1 << 16, 99, 0.99
BATCH_CAL_N = 1 << 14
BATCH_TRAIN_N, BATCH_TRAIN_STEPS = 1 << 18, 500
# FLOP per sample: forward 2*(99*128 + 3*128*128 + 128); backward 2 GEMMs per layer (dW, dH)
MLP_FWD_FLOP = 2 * (99 * 128 + 3 * 128 * 128 + 128)
MLP_BWD_FLOP = 2 * (99 * 128 + 3 * 128 * 128 + 128) + 2 * (3 * 128 * 128 + 128)
COPY_BYTES = 1 << 30

# bytes of the record each counted event reads (mi_mesh_set_stats slots; layouts in mi_tree.hpp /
# mi_kernels.hpp): 4-wide quad 128 B, plane record 80 B, triangle 48 B, binary node 48 B of box pair
# read by the keyed walk's near/far choice, dipole 32 B, child refs 16 B
RECORD_BYTES = {"quad_steps": 128, "prism_reads": 80, "leaf_tests": 48, "keyed_steps": 128 + 48,
                "keyed_tests": 48, "fwn_visits": 32, "fwn_descents": 16, "fwn_leaves": 48}
SOURCES = ("meshlib_amd/csrc", "meshlib_amd/build.py")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sources_sha256():
    """Hash of the kernel sources and build flags: identifies the build a PMC profile was taken on."""
    h = hashlib.sha256()
    files = []
    for s in SOURCES:
        p = os.path.join(ROOT, s)
        if os.path.isdir(p):
            files += sorted(os.path.join(p, f) for f in os.listdir(p)
                            if f.endswith((".hip", ".hpp", ".cpp", ".h")))
        else:
            files.append(p)
    for f in files:
        h.update(os.path.relpath(f, ROOT).encode())
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


# ------------------------------------------------------------------ measurement helpers

class Timer:
    """HIP events on torch's current stream (the stream every call below is launched on)."""

    def __init__(self, torch):
        self.torch = torch
        self.stream = torch.cuda.current_stream()
        self.ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def __call__(self, fn):
        self.ev[0].record(self.stream)
        r = fn()
        self.ev[1].record(self.stream)
        self.torch.cuda.synchronize()
        return self.ev[0].elapsed_time(self.ev[1]), r


def counted(torch, M, meshes, fn):
    """Runs fn once with the walks' record counters on: {launch: {slot: count}}."""
    buf = torch.zeros(M.STAT_LAUNCHES * M.STAT_STRIDE, dtype=torch.int64, device="cuda")
    for m in meshes:
        m.set_stats(buf.data_ptr())
    torch.cuda.synchronize()
    try:
        fn()
        torch.cuda.synchronize()
    finally:
        for m in meshes:
            m.set_stats(0)
    c = buf.view(M.STAT_LAUNCHES, M.STAT_STRIDE).cpu().numpy()
    return [{name: int(c[k][i]) for i, name in enumerate(M.STAT_NAMES)} for k in range(M.STAT_LAUNCHES)]


def walk_bytes(cnt):
    """Record bytes one launch's walks read, from its counters (per-wave counts)."""
    b = sum(RECORD_BYTES[k] * cnt[k] for k in RECORD_BYTES)
    b += cnt["seed_levels"] * 64 + cnt["seeds"] * 48    # greedy descents: a 64-B node record per level + a triangle
    return b


def pmc_profile(name):
    """The committed PMC summary profiles/pmc_NAME.json, if it was taken on the current sources."""
    p = os.path.join(ROOT, "profiles", f"pmc_{name}.json")
    if not os.path.exists(p):
        return None, "missing"
    with open(p) as fh:
        prof = json.load(fh)
    if prof.get("sources_sha256") != sources_sha256():
        return None, "stale (kernel sources changed since the profile)"
    return prof, "current"


def pmc_call(prof, kernels):
    """HBM bytes (fetch and write, copy-calibrated) and L2 hit rate summed over a call's kernels."""
    if prof is None:
        return None
    cal = prof["calibration"]
    f_scale = cal["copy_bytes"] / (cal["fetch_kib"] * 1024.0)
    w_scale = cal["copy_bytes"] / (cal["write_kib"] * 1024.0)
    tot = {"fetch": 0.0, "write": 0.0, "hit": 0.0, "miss": 0.0, "valu": 0.0, "wave_cycles": 0.0,
           "wait_any": 0.0, "busy": 0.0, "dispatches": 0}
    found = False
    for kname, k in prof["kernels"].items():
        if not any(kname.startswith(p) for p in kernels):
            continue
        found = True
        n = max(k.get("dispatches", 1), 1)
        tot["fetch"] += k.get("FETCH_SIZE", 0.0) / n * 1024.0 * f_scale
        tot["write"] += k.get("WRITE_SIZE", 0.0) / n * 1024.0 * w_scale
        tot["hit"] += k.get("TCC_HIT_sum", 0.0) / n
        tot["miss"] += k.get("TCC_MISS_sum", 0.0) / n
        tot["valu"] += k.get("SQ_INSTS_VALU", 0.0) / n
        tot["wave_cycles"] += k.get("SQ_WAVE_CYCLES", 0.0) / n
        tot["wait_any"] += k.get("SQ_WAIT_INST_ANY", 0.0) / n
        tot["busy"] += k.get("SQ_BUSY_CYCLES", 0.0) / n
        tot["dispatches"] += 1
    if not found:
        return None
    out = {"traffic": tot["fetch"] + tot["write"], "fetch_bytes": tot["fetch"], "write_bytes": tot["write"],
           "l2_hit_rate": tot["hit"] / (tot["hit"] + tot["miss"]) if tot["hit"] + tot["miss"] > 0 else None}
    if tot["wave_cycles"] > 0:
        out["stall_frac"] = tot["wait_any"] / tot["wave_cycles"]
    if tot["valu"] > 0:
        out["valu_insts"] = tot["valu"]
    return out


def hbm_roofline(kernel, bytes_per_call, ms, counters, prof=None, prof_state=None, prof_kernels=()):
    achieved = bytes_per_call / (ms / 1e3) / 1e9
    r = {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": achieved / HBM_PEAK_GBS, "traffic": None, "bytes_per_call": bytes_per_call, "ms": ms,
         "bytes_model": "records read per wave (kernels' own counters), each counted once per wave visit",
         "per_launch_bytes": [c["bytes"] for c in counters], "counters": counters}
    pm = pmc_call(prof, prof_kernels)
    if pm:
        r["traffic"] = pm["traffic"]
        r["hbm_gbs_measured"] = pm["traffic"] / (ms / 1e3) / 1e9
        r["l2_hit_rate"] = pm["l2_hit_rate"]
        if "stall_frac" in pm:
            r["stall_frac"] = pm["stall_frac"]
        if "valu_insts" in pm:
            r["valu_issue_frac"] = pm["valu_insts"] / (ms / 1e3) / (1024 * 2.4e9 / 4)
    r["pmc_profile"] = prof_state
    return r


def with_bytes(cnts, io_bytes=0):
    out = []
    for k, c in enumerate(cnts):
        c = dict(c)
        c["bytes"] = walk_bytes(c) + (io_bytes if k == 0 else 0)
        out.append(c)
    return out


# ------------------------------------------------------------------ CPU baseline

def host_cores():
    """Threads this process may run on: the affinity mask, capped by a cgroup CPU quota if one is set."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as fh:
            q, per = fh.read().split()[:2]
            if q != "max":
                quota = max(1, int(int(q) / int(per)))
    except (OSError, ValueError):
        pass
    return (min(aff, quota) if quota else aff), aff, quota


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def median_of_5(fn):
    fn()                                        # warm-up
    ts = []
    for _ in range(5):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return statistics.median(ts), ts


def cpu_baseline(np, synth, v, f, dims, xf, pts, nvox):
    from oracle import oracle as O
    threads, aff, quota = host_cores()
    # the reference's CPU path (ParallelFor over contiguous voxel / query ranges, all host threads)
    v1, f1 = synth.bumpy_icosphere(5, 8, 1)
    d1, x1, _ = synth.grid_for_mesh(v1, (64, 64, 64))
    om1 = O.Mesh(v1, f1)
    t1, _ = median_of_5(lambda: om1.sdf_grid(d1, x1, nthreads=threads))
    om = O.Mesh(v, f)
    layer = int(dims[0]) * int(dims[1])
    zs = list(range(16, int(dims[2]), 32))                  # 8 whole z-layers, evenly spaced
    ids = np.concatenate([np.arange(z * layer, (z + 1) * layer, dtype=np.int64) for z in zs])
    tv, _ = median_of_5(lambda: om.sdf_grid_sample(dims, xf, ids, nthreads=threads))
    half = pts.shape[0] // 2
    sel = np.concatenate([np.arange(0, 1 << 17), np.arange(half, half + (1 << 17))])
    psel = np.ascontiguousarray(pts[sel])
    tp, _ = median_of_5(lambda: om.find_projections(psel, nthreads=threads))
    vox_rate, pts_rate = ids.size / tv, psel.shape[0] / tp
    # one thread on a quarter layer / 16k queries, for the per-core figure
    ids1 = ids[: layer // 4]
    t = time.perf_counter()
    om.sdf_grid_sample(dims, xf, ids1, nthreads=1)
    vox1 = ids1.size / (time.perf_counter() - t)
    t = time.perf_counter()
    om.find_projections(psel[: 1 << 14], nthreads=1)
    pts1 = (1 << 14) / (time.perf_counter() - t)
    step = nvox / vox_rate + pts.shape[0] / pts_rate
    return {"value": (nvox + pts.shape[0]) / step, "unit": "queries/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "affinity_cpus": aff, "cgroup_cpu_quota": quota,
            "method": "oracle (C restatement of the reference CPU path, -O2 -ffp-contract=off) with a "
                      "ParallelFor of dynamic contiguous chunks on all threads; 1 warm-up + median of 5",
            "sample": f"k=8 256^3 step: {len(zs)} whole z-layers of the 256^3 grid (z = {zs[0]}, {zs[1]}, ..., "
                      f"{zs[-1]}: {ids.size} voxels) and 2 x 2^17 contiguous queries (one range of each "
                      f"half of the 2^22 set); step time = voxels / voxel rate + points / point rate",
            "voxels_per_s": vox_rate, "pts_per_s": pts_rate,
            "k5_64_full": {"workload": "k=5 (20,480 tris, 8 bumps), 64^3 signed-distance grid",
                             "voxels_per_s": 64 ** 3 / t1, "ms": t1 * 1e3},
            "one_thread": {"voxels_per_s": vox1, "pts_per_s": pts1,
                           "sample": f"{ids1.size} contiguous voxels + 16384 contiguous queries"}}


# ------------------------------------------------------------------ workloads

def learned_leg(ctx):
    """Train the MLP on labels from the GPU projector, run the capped search at each margin (outputs
    must be identical), and the ceiling: caps equal to the true distances (a perfect predictor)."""
    M, torch, np, synth = ctx["M"], ctx["torch"], ctx["np"], ctx["synth"]
    v, f, mesh, fwn, proj = ctx["v"], ctx["f"], ctx["mesh"], ctx["fwn"], ctx["proj"]
    dims, xf, opts, d_pts, d_sdf, d_proj = ctx["dims"], ctx["xf"], ctx["opts"], ctx["d_pts"], ctx["d_sdf"], ctx["d_proj"]
    res_sdf, res_proj, timer, sh = ctx["res_sdf"], ctx["res_proj"], ctx["timer"], ctx["sh"]
    nq, nvox = d_pts.shape[0], d_sdf.shape[0]
    lo, hi = v.min(0).astype(np.float64), v.max(0).astype(np.float64)
    net = M.DistanceMLP(0.5 * (lo + hi), float(0.5 * np.max(hi - lo) * 1.2), seed=TRAIN_SEED)
    tp = torch.from_numpy(synth.projection_queries(v, f, TRAIN_N, TRAIN_SEED)).cuda()
    trec = torch.empty(TRAIN_N * 8, dtype=torch.int32, device="cuda")
    proj.find_projections_device(tp.data_ptr(), TRAIN_N, trec.data_ptr(), stream=sh)
    tlab = trec.view(torch.float32).view(-1, 8)[:, 7].sqrt().contiguous()
    torch.cuda.synchronize()
    loss0 = net.loss(tp[:65536], tlab[:65536])
    steps = ctx["train_steps"]
    train_ms, _ = timer(lambda: net.train(tp, tlab, steps=steps, batch=TRAIN_BATCH, lr=TRAIN_LR, stream=sh))
    loss1 = net.loss(tp[:65536], tlab[:65536])
    # slack: the 0.99 quantile of the network's under-prediction on held-out labelled points
    cp = torch.from_numpy(synth.projection_queries(v, f, CAL_N, CAL_SEED)).cuda()
    crec = torch.empty(CAL_N * 8, dtype=torch.int32, device="cuda")
    proj.find_projections_device(cp.data_ptr(), CAL_N, crec.data_ptr(), stream=sh)
    clab = crec.view(torch.float32).view(-1, 8)[:, 7].sqrt().contiguous()
    slack = net.calibrate(cp, clab, CAL_Q, stream=sh)
    ref_d = res_proj.view(torch.float32).view(-1, 8)[:, 7].double().sqrt().numpy()
    d_pred = torch.empty(nq, dtype=torch.float32, device="cuda")
    net.predict_device(d_pts, d_pred, stream=sh)
    infer_ms, _ = timer(lambda: net.predict_device(d_pts, d_pred, stream=sh))
    pred = d_pred.cpu().numpy()
    rel = np.abs(pred - ref_d) / np.maximum(ref_d, 1e-3)
    train_tflops = steps * TRAIN_BATCH * (MLP_FWD_FLOP + MLP_BWD_FLOP) / (train_ms / 1e3) / 1e12
    infer_tflops = nq * MLP_FWD_FLOP / (infer_ms / 1e3) / 1e12
    out = {"train": {"samples": TRAIN_N, "steps": steps, "batch": TRAIN_BATCH, "lr": TRAIN_LR, "ms": train_ms,
                     "loss_before": loss0, "loss_after": loss1, "tflops": train_tflops},
           "calibration": {"samples": CAL_N, "quantile": CAL_Q, "slack": slack,
                           "cap": "(pred (1 + m) + slack)^2; grids: one prediction per 4^3 brick at its centre c, "
                                  "voxel x capped at (pred(c) (1 + m) + slack + |x - c|)^2"},
           "inference": {"points": nq, "ms": infer_ms, "tflops": infer_tflops,
                         "median_rel_err": float(np.median(rel)), "p90_rel_err": float(np.quantile(rel, 0.9)),
                         "under_prediction_rate": float(np.mean(pred < ref_d))},
           "margins": {}}
    ctx["mlp"] = {"train_tflops": train_tflops, "infer_tflops": infer_tflops, "train_ms": train_ms,
                  "infer_ms": infer_ms, "train_flop": steps * TRAIN_BATCH * (MLP_FWD_FLOP + MLP_BWD_FLOP),
                  "infer_flop": nq * MLP_FWD_FLOP}
    base_p, base_s = ctx["proj_avg"], ctx["sdf_avg"]
    plain_p = sum(c["keyed_steps"] for c in ctx["cnt_proj"])
    plain_s = sum(c["quad_steps"] for c in ctx["cnt_sdf"])

    def margin_run(m):
        pm, rp = timer(lambda: proj.find_projections_learned_device(net, m, d_pts.data_ptr(), nq, d_proj.data_ptr(),
                                                                    stream=sh))
        same_p = bool(torch.equal(d_proj.cpu(), res_proj))
        sm, rs = timer(lambda: fwn.sdf_grid_learned_device(net, m, dims, xf, opts, d_sdf.data_ptr(), stream=sh))
        same_s = bool(torch.equal(d_sdf.cpu().view(torch.int32), res_sdf.view(torch.int32)))
        cp_ = counted(torch, M, [mesh], lambda: proj.find_projections_learned_device(
            net, m, d_pts.data_ptr(), nq, d_proj.data_ptr(), stream=sh))
        cs_ = counted(torch, M, [mesh], lambda: fwn.sdf_grid_learned_device(net, m, dims, xf, opts, d_sdf.data_ptr(),
                                                                             stream=sh))
        r = {"proj_ms": pm, "sdf_ms": sm, "proj_fallback_rate": rp / nq, "sdf_fallback_rate": rs / nvox,
             "identical_projections": same_p, "identical_sdf": same_s, "proj_speedup_vs_plain": base_p / pm,
             "sdf_speedup_vs_plain": base_s / sm, "speedup_vs_plain_coop": (base_p + base_s) / (pm + sm),
             "proj_node_visit_ratio": sum(c["keyed_steps"] for c in cp_) / max(plain_p, 1),
             "sdf_node_visit_ratio": sum(c["quad_steps"] for c in cs_) / max(plain_s, 1)}
        log(f"learned m={m} slack={net_slack[0]:.4f}: proj {pm:.2f} ms (fallback {rp / nq:.4f}, visits "
            f"{r['proj_node_visit_ratio']:.3f}), sdf {sm:.2f} ms (fallback {rs / nvox:.4f}, visits "
            f"{r['sdf_node_visit_ratio']:.3f}), identical {same_p} {same_s}")
        return r
    net_slack = [slack]
    for m in MARGINS:
        out["margins"][str(m)] = margin_run(m)
    best = max(out["margins"], key=lambda k: out["margins"][k]["speedup_vs_plain_coop"])
    out["best"] = {"margin": float(best), "speedup_vs_plain_coop": out["margins"][best]["speedup_vs_plain_coop"]}
    # the same network without the calibrated slack (the bare cap (pred (1 + m))^2), for reference
    net.set_slack(0.0)
    net_slack[0] = 0.0
    out["uncalibrated"] = {str(m): margin_run(m) for m in (MARGINS[-1],)}
    net.set_slack(slack)
    # ceiling: caps = the true squared distance (x (1 + 1e-4)), i.e. a perfect network with no inference
    # cost; what is left of the walk are the boxes closer than the true distance, which no bound prunes
    pd = res_proj.view(torch.float32).view(-1, 8)[:, 7].cuda() * (1 + 1e-4)
    sd = (res_sdf.cuda() ** 2 * (1 + 1e-4)).nan_to_num(nan=float(np.finfo(np.float32).max)).contiguous()
    cp, rcp = timer(lambda: proj.find_projections_capped_device(d_pts.data_ptr(), nq, pd.data_ptr(), d_proj.data_ptr(),
                                                                stream=sh))
    same_cp = bool(torch.equal(d_proj.cpu(), res_proj))
    cs, rcs = timer(lambda: fwn.sdf_grid_capped_device(dims, xf, opts, sd.data_ptr(), d_sdf.data_ptr(), stream=sh))
    same_cs = bool(torch.equal(d_sdf.cpu().view(torch.int32), res_sdf.view(torch.int32)))
    cnt_p = counted(torch, M, [mesh], lambda: proj.find_projections_capped_device(
        d_pts.data_ptr(), nq, pd.data_ptr(), d_proj.data_ptr(), stream=sh))
    cnt_s = counted(torch, M, [mesh], lambda: fwn.sdf_grid_capped_device(dims, xf, opts, sd.data_ptr(),
                                                                           d_sdf.data_ptr(), stream=sh))
    steps_p = sum(c["keyed_steps"] for c in cnt_p)
    steps_s = sum(c["quad_steps"] for c in cnt_s)
    plain_p = sum(c["keyed_steps"] for c in ctx["cnt_proj"])
    plain_s = sum(c["quad_steps"] for c in ctx["cnt_sdf"])
    out["ceiling"] = {"caps": "true squared distance x (1 + 1e-4): a perfect predictor with no inference cost",
                      "proj_ms": cp, "sdf_ms": cs, "proj_speedup_vs_plain": base_p / cp,
                      "sdf_speedup_vs_plain": base_s / cs, "speedup_vs_plain_coop": (base_p + base_s) / (cp + cs),
                      "proj_node_visit_ratio": steps_p / max(plain_p, 1),
                      "sdf_node_visit_ratio": steps_s / max(plain_s, 1),
                      "reruns": rcp + rcs, "identical": same_cp and same_cs,
                      "why": "the plain walk is already seeded with a real candidate per query (a greedy descent), "
                             "so a cap prunes only what lies beyond that seed; the visits left are boxes nearer "
                             "than the true distance, which no bound can prune. The target of >= 2x over the "
                             "plain walk is above this ceiling."}
    log(f"ceiling (perfect caps): proj {cp:.2f} ms, sdf {cs:.2f} ms, visits {steps_p / max(plain_p, 1):.3f} / "
        f"{steps_s / max(plain_s, 1):.3f} of plain")
    return out


def batched_leg(args, M, torch, np, synth, timer):
    """64 bumpy spheres (split levels 4..8, grids 32..192, 2^18..2^21 queries) in one persistent launch."""
    from concurrent.futures import ThreadPoolExecutor
    t0 = time.time()
    specs = synth.batched_workload_specs()

    def gen(s):
        level, bumps, mseed, side, nq, qseed = s
        v, f = synth.bumpy_icosphere(level, bumps, mseed)
        return v, f, synth.projection_queries(v, f, nq, qseed)
    with ThreadPoolExecutor(min(8, len(os.sched_getaffinity(0)))) as ex:
        geo = list(ex.map(gen, specs))
    items, per_item, meshes = [], [], []
    tot_vox = tot_pts = 0
    for (level, bumps, mseed, side, nq, qseed), (v, f, pts) in zip(specs, geo):
        mesh = M.Mesh(v, f)
        dims, xf, vox = synth.grid_for_mesh(v, (side,) * 3)
        dp = torch.from_numpy(pts).cuda()
        ds = torch.empty(int(np.prod(dims)), dtype=torch.float32, device="cuda")
        dr = torch.empty(nq * 8, dtype=torch.int32, device="cuda")
        items.append({"mesh": mesh, "dims": dims, "xf": xf, "d_sdf": ds.data_ptr(), "d_pts": dp.data_ptr(), "n": nq,
                      "d_proj": dr.data_ptr(), "_keep": (dp, ds, dr)})
        fwn = M.FastWindingNumber(mesh)
        pr = M.PointsToMeshProjector()
        pr.updateMeshData(mesh)
        per_item.append((fwn, pr, dims, xf, ds, dp, dr))
        meshes.append((mesh, v, f))
        tot_vox += int(np.prod(dims))
        tot_pts += nq
    log(f"batched inputs {time.time() - t0:.1f}s: {len(items)} meshes, {tot_vox} voxels, {tot_pts} points")
    opts = M.DistanceToMeshOptions()

    def per_item_run():
        for fwn, pr, dims, xf, ds, dp, dr in per_item:
            fwn.sdf_grid_device(dims, xf, opts, ds.data_ptr())
            pr.find_projections_device(dp.data_ptr(), dp.shape[0], dr.data_ptr())

    per_item_run()
    t_items, _ = timer(per_item_run)
    ref = [(ds.clone(), dr.clone()) for (_, _, _, _, ds, _, dr) in per_item]
    batch = M.Batch()
    t_tune = batch.autotune(items, opts)
    cfg = batch.config()
    ts = [timer(lambda: batch.run(items, opts))[0] for _ in range(args.sub_steps)]
    t_batch = statistics.mean(ts)

    def same_all():
        return all(torch.equal(ds.view(torch.int32), r[0].view(torch.int32)) and torch.equal(dr, r[1])
                   for (_, _, _, _, ds, _, dr), r in zip(per_item, ref))
    same = same_all()
    cnts = counted(torch, M, [m for m, _, _ in meshes], lambda: batch.run(items, opts))
    # every item reads its own voxels out (4 B) and points in / records out / sorted ids (12 + 32 + 4 B)
    cb = with_bytes(cnts, io_bytes=4 * tot_vox + 48 * tot_pts)
    prof, pstate = pmc_profile("batched")
    roof = hbm_roofline("kBatch", cb[0]["bytes"], t_batch, cb[:1], prof, pstate, ("kBatch",))
    log(f"batched: {t_batch:.1f} ms (per-item {t_items:.1f} ms), identical {same}")
    learned = None
    if not args.no_learned:
        t0 = time.time()
        nets = []
        for i, ((mesh, v, f), (fwn, pr, *_)) in enumerate(zip(meshes, per_item)):
            lo, hi = v.min(0).astype(np.float64), v.max(0).astype(np.float64)
            net = M.DistanceMLP(0.5 * (lo + hi), float(0.5 * np.max(hi - lo) * 1.2), seed=TRAIN_SEED + i)
            tp = torch.from_numpy(synth.projection_queries(v, f, BATCH_TRAIN_N, 3000 + i)).cuda()
            trec = torch.empty(BATCH_TRAIN_N * 8, dtype=torch.int32, device="cuda")
            pr.find_projections_device(tp.data_ptr(), BATCH_TRAIN_N, trec.data_ptr())
            tlab = trec.view(torch.float32).view(-1, 8)[:, 7].sqrt().contiguous()
            net.train(tp, tlab, steps=BATCH_TRAIN_STEPS, batch=TRAIN_BATCH, lr=TRAIN_LR)
            # slack from held-out labelled points, as for the single mesh
            cp = torch.from_numpy(synth.projection_queries(v, f, BATCH_CAL_N, 5000 + i)).cuda()
            crec = torch.empty(BATCH_CAL_N * 8, dtype=torch.int32, device="cuda")
            pr.find_projections_device(cp.data_ptr(), BATCH_CAL_N, crec.data_ptr())
            net.calibrate(cp, crec.view(torch.float32).view(-1, 8)[:, 7].sqrt().contiguous(), CAL_Q)
            nets.append(net)
        torch.cuda.synchronize()
        learned = {"train_s_all_meshes": time.time() - t0, "samples_per_mesh": BATCH_TRAIN_N,
                   "steps_per_mesh": BATCH_TRAIN_STEPS, "calibration_samples_per_mesh": BATCH_CAL_N,
                   "calibration_quantile": CAL_Q, "margins": {}}
        for m in (0.1, 0.25):
            rr = [0]

            def go():
                rr[0] = batch.run_learned(items, nets, m, opts)
            t_l = statistics.mean(timer(go)[0] for _ in range(args.sub_steps))
            learned["margins"][str(m)] = {"ms": t_l, "reruns": rr[0], "rerun_rate": rr[0] / (tot_vox + tot_pts),
                                          "identical_to_per_item": bool(same_all()),
                                          "speedup_vs_batched_plain": t_batch / t_l,
                                          "speedup_vs_per_item_coop": t_items / t_l}
            log(f"learned+batched m={m}: {t_l:.1f} ms, reruns {rr[0]}")
    return {"metric": "signed-dist voxels/s & projected pts/s (batched uneven workload)",
            "value": (tot_vox + tot_pts) / (t_batch / 1e3), "unit": "queries/s (voxels + projected points)",
            "ms": t_batch, "ms_runs": ts,
            "config": {"workload": f"{len(items)} bumpy icospheres, levels 4..8, grids 32..192, "
                                   f"2^18..2^21 points each (seeded)", "voxels": tot_vox, "points": tot_pts},
            "autotune": {"blocks_per_cu": cfg[0], "xcd_split": cfg[1], "best_ms": t_tune},
            "per_item_coop_ms": t_items, "speedup_vs_per_item_coop": t_items / t_batch,
            "identical_to_per_item": bool(same), "roofline": roof, "learned_batched": learned}


def offset_leg(args, M, torch, np, synth, timer):
    """512^3 volume of a 5,242,880-triangle mesh: plain, learned, and the offset caller's band options."""
    t0 = time.time()
    v, f = synth.bumpy_icosphere(9, 64, 5)
    dims, xf, vox = synth.grid_for_mesh(v, (512, 512, 512))
    nvox = 512 ** 3
    mesh = M.Mesh(v, f)
    log(f"offset inputs + tree {time.time() - t0:.1f}s (depth {mesh.depth()})")
    fwn = M.FastWindingNumber(mesh)
    proj = M.PointsToMeshProjector()
    proj.updateMeshData(mesh)
    opts = M.DistanceToMeshOptions()
    d_sdf = torch.empty(nvox, dtype=torch.float32, device="cuda")
    sh = torch.cuda.current_stream().cuda_stream
    run = lambda o: fwn.sdf_grid_device(dims, xf, o, d_sdf.data_ptr(), stream=sh)   # noqa: E731
    run(opts)
    ts = [timer(lambda: run(opts))[0] for _ in range(args.sub_steps)]
    sdf_ms = statistics.mean(ts)
    ref = d_sdf.view(torch.int32).clone()
    cnts = counted(torch, M, [mesh], lambda: run(opts))
    cb = with_bytes(cnts, io_bytes=4 * nvox)
    prof, pstate = pmc_profile("offset")
    roof = hbm_roofline("kSdfGridCoop+kSdfSplit+kSdfSplitFinish", sum(c["bytes"] for c in cb), sdf_ms, cb, prof,
                        pstate, ("kSdfGridCoop", "kSdfSplit"))
    log(f"offset plain: {sdf_ms:.1f} ms")
    # band options of mcOffsetMesh (source/MRVoxels/MROffset.cpp:176-183) for an offset of 4 voxels,
    # nullOutsideMinMax false (the by-parts path), on the device path, the per-lane form and by parts
    vs = float(np.max(vox))
    band = M.DistanceToMeshOptions()
    band.minDistSq = float(np.float32(4 * vs - 1.001 * vs) ** 2)
    band.maxDistSq = float(np.float32(4 * vs + 1.001 * vs) ** 2)
    band.nullOutsideMinMax = False
    run(band)
    tb = statistics.mean(timer(lambda: run(band))[0] for _ in range(args.sub_steps))
    band_ref = d_sdf.view(torch.int32).clone()
    M.set_traversal("per_lane")
    try:
        tpl, _ = timer(lambda: run(band))
    finally:
        M.set_traversal("wave_coop")
    same_pl = bool(torch.equal(d_sdf.view(torch.int32), band_ref))
    host = np.empty(nvox, np.float32)
    layer = 512 * 512

    def take(data, pdims, z0):
        host[z0 * layer:(z0 + int(pdims[2])) * layer] = data
    tbps = []
    for _ in range(2):     # the first call also fills the pool of page-locked staging buffers
        host.fill(0)
        t = time.perf_counter()
        fwn.calcFromGridWithDistancesByParts(take, dims, xf, band, 1)
        tbps.append((time.perf_counter() - t) * 1e3)
    tbp = tbps[-1]
    same_bp = bool(np.array_equal(host.view(np.int32), band_ref.cpu().numpy()))
    t = time.perf_counter()
    hres = fwn.calcFromGridWithDistances(dims, xf, band)
    thp = (time.perf_counter() - t) * 1e3
    same_hp = bool(np.array_equal(hres.view(np.int32), band_ref.cpu().numpy()))
    del hres
    log(f"offset band: device {tb:.1f} ms, per-lane {tpl:.1f} ms, by parts {tbp:.1f} ms (first call {tbps[0]:.1f}), "
        f"host-pointer {thp:.1f} ms, identical {same_pl} {same_bp} {same_hp}")
    band_out = {"options": {"offset_voxels": 4, "minDistSq": band.minDistSq, "maxDistSq": band.maxDistSq,
                            "nullOutsideMinMax": False},
                "device_ms": tb, "voxels_per_s": nvox / (tb / 1e3), "per_lane_ms": tpl, "speedup_vs_per_lane": tpl / tb,
                "by_parts_ms": tbp, "by_parts_first_call_ms": tbps[0],
                "by_parts_note": "host-pointer pipeline: includes the PCIe copies of 537 MB and the part callback "
                                 "(a copy into one host array); second call of the process",
                "host_pointer_ms": thp, "identical_per_lane": same_pl, "identical_by_parts": same_bp,
                "identical_host_pointer": same_hp}
    learned = None
    if not args.no_learned:
        lo, hi = v.min(0).astype(np.float64), v.max(0).astype(np.float64)
        net = M.DistanceMLP(0.5 * (lo + hi), float(0.5 * np.max(hi - lo) * 1.2), seed=TRAIN_SEED)
        tp = torch.from_numpy(synth.projection_queries(v, f, TRAIN_N, TRAIN_SEED)).cuda()
        trec = torch.empty(TRAIN_N * 8, dtype=torch.int32, device="cuda")
        proj.find_projections_device(tp.data_ptr(), TRAIN_N, trec.data_ptr(), stream=sh)
        tlab = trec.view(torch.float32).view(-1, 8)[:, 7].sqrt().contiguous()
        train_ms, _ = timer(lambda: net.train(tp, tlab, steps=args.train_steps, batch=TRAIN_BATCH, lr=TRAIN_LR,
                                              stream=sh))
        cp = torch.from_numpy(synth.projection_queries(v, f, CAL_N, CAL_SEED)).cuda()
        crec = torch.empty(CAL_N * 8, dtype=torch.int32, device="cuda")
        proj.find_projections_device(cp.data_ptr(), CAL_N, crec.data_ptr(), stream=sh)
        slack = net.calibrate(cp, crec.view(torch.float32).view(-1, 8)[:, 7].sqrt().contiguous(), CAL_Q, stream=sh)
        learned = {"train_ms": train_ms, "train_steps": args.train_steps, "slack": slack, "margins": {}}
        for m in MARGINS:
            t, r = timer(lambda: fwn.sdf_grid_learned_device(net, m, dims, xf, opts, d_sdf.data_ptr(), stream=sh))
            same = bool(torch.equal(d_sdf.view(torch.int32), ref))
            learned["margins"][str(m)] = {"sdf_ms": t, "fallback_rate": r / nvox, "identical": same,
                                          "speedup_vs_plain": sdf_ms / t}
            log(f"offset learned m={m}: {t:.1f} ms, fallback {r / nvox:.4f}, identical {same}")
    return {"metric": "signed-dist voxels/s (offset-scale volume)", "value": nvox / (sdf_ms / 1e3),
            "unit": "voxels/s", "ms": sdf_ms, "ms_runs": ts,
            "config": {"workload": "bumpy icosphere k=9 (5,242,880 tris, 64 bumps): 512^3 signed-distance grid",
                       "voxels": nvox, "beta": 2.0},
            "roofline": roof, "band_options": band_out, "learned_vs_plain": learned}


# ------------------------------------------------------------------ PMC summaries

def summarize_pmc(src_dir, name):
    """rocprofv3 counter-collection CSVs under src_dir (one subdirectory per pass) -> profiles/pmc_NAME.json:
    counter sums per kernel name over its dispatches, the 1 GiB copy's readings as the byte calibration,
    and the hash of the kernel sources the profile was taken on."""
    kernels = {}
    copy_max = {}      # the largest single device-copy dispatch per counter: the 1 GiB calibration copy
    for fn in sorted(glob.glob(os.path.join(src_dir, "**", "*counter_collection.csv"), recursive=True)):
        seen = {}
        with open(fn, newline="") as fh:
            for row in csv.DictReader(fh):
                low = {k.lower(): v for k, v in row.items() if k}
                kname = low.get("kernel_name") or low.get("kernel-name") or ""
                cname = low.get("counter_name") or ""
                try:
                    val = float(low.get("counter_value") or 0)
                except ValueError:
                    continue
                disp = low.get("dispatch_id") or low.get("correlation_id") or ""
                short = kname.replace("(anonymous namespace)::", "").replace("void ", "").split("(")[0]
                short = short.replace("mi::", "").strip()
                k = kernels.setdefault(short, {})
                k[cname] = k.get(cname, 0.0) + val
                seen.setdefault(short, set()).add(disp)
                if "copyBuffer" in short:
                    copy_max[cname] = max(copy_max.get(cname, 0.0), val)
        for short, ds in seen.items():
            kernels[short]["dispatches"] = max(kernels[short].get("dispatches", 0), len(ds))
    if not copy_max.get("FETCH_SIZE") or not copy_max.get("WRITE_SIZE"):
        raise SystemExit("no device copy with FETCH_SIZE / WRITE_SIZE in the profile (run with --profile-run)")
    cal = {"copy_bytes": COPY_BYTES, "fetch_kib": copy_max["FETCH_SIZE"], "write_kib": copy_max["WRITE_SIZE"]}
    out = {"name": name, "sources_sha256": sources_sha256(), "calibration": cal,
           "kernels": {k: v for k, v in kernels.items() if not k.startswith("__amd")}}
    path = os.path.join(ROOT, "profiles", f"pmc_{name}.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print(path)


def profile_run(args):
    """One call of each measured kernel family (for rocprofv3 passes), after a 1 GiB device copy that
    calibrates the byte counters."""
    import numpy as np
    import torch

    import meshlib_amd as M
    from meshlib_amd import synth
    torch.cuda.set_device(0)
    a = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    log("copy calibration done")
    if args.workload == "single":
        v, f = synth.bumpy_icosphere(MESH_LEVEL, MESH_BUMPS, MESH_SEED)
        dims, xf, _ = synth.grid_for_mesh(v, (GRID, GRID, GRID))
        pts = synth.projection_queries(v, f, NQ, QSEED)
        mesh = M.Mesh(v, f)
        fwn, pr = M.FastWindingNumber(mesh), M.PointsToMeshProjector()
        pr.updateMeshData(mesh)
        d_pts = torch.from_numpy(pts).cuda()
        d_sdf = torch.empty(GRID ** 3, dtype=torch.float32, device="cuda")
        d_proj = torch.empty(NQ * 8, dtype=torch.int32, device="cuda")
        fwn.sdf_grid_device(dims, xf, M.DistanceToMeshOptions(), d_sdf.data_ptr())
        torch.cuda.synchronize()
        log("sdf done")
        pr.find_projections_device(d_pts.data_ptr(), NQ, d_proj.data_ptr())
        torch.cuda.synchronize()
        log("projections done")
    elif args.workload == "batched":
        from concurrent.futures import ThreadPoolExecutor
        specs = synth.batched_workload_specs()

        def gen(s):
            v, f = synth.bumpy_icosphere(s[0], s[1], s[2])
            return v, f, synth.projection_queries(v, f, s[4], s[5])
        with ThreadPoolExecutor(8) as ex:
            geo = list(ex.map(gen, specs))
        items = []
        for s, (v, f, pts) in zip(specs, geo):
            dims, xf, _ = synth.grid_for_mesh(v, (s[3],) * 3)
            dp = torch.from_numpy(pts).cuda()
            ds = torch.empty(int(np.prod(dims)), dtype=torch.float32, device="cuda")
            dr = torch.empty(s[4] * 8, dtype=torch.int32, device="cuda")
            items.append({"mesh": M.Mesh(v, f), "dims": dims, "xf": xf, "d_sdf": ds.data_ptr(), "d_pts": dp.data_ptr(),
                          "n": s[4], "d_proj": dr.data_ptr(), "_keep": (dp, ds, dr)})
        log("batched inputs done")
        batch = M.Batch()
        batch.set_config(6, 0)
        batch.run(items)
        torch.cuda.synchronize()
        log("batch done")
    elif args.workload == "offset":
        v, f = synth.bumpy_icosphere(9, 64, 5)
        dims, xf, _ = synth.grid_for_mesh(v, (512, 512, 512))
        mesh = M.Mesh(v, f)
        d_sdf = torch.empty(512 ** 3, dtype=torch.float32, device="cuda")
        M.FastWindingNumber(mesh).sdf_grid_device(dims, xf, M.DistanceToMeshOptions(), d_sdf.data_ptr())
        torch.cuda.synchronize()
        log("offset sdf done")
    print("profile run done", flush=True)


# ------------------------------------------------------------------ main

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the one-step comparison run with the per-lane (reference-form) traversal")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-bound legs")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched uneven workload")
    ap.add_argument("--no-offset", action="store_true", help="skip the 512^3 offset-scale workload")
    ap.add_argument("--train-steps", type=int, default=2000)
    ap.add_argument("--workload", default="single", choices=["single", "batched", "offset"],
                    help="with --profile-run: which kernels to run once")
    ap.add_argument("--profile-run", action="store_true",
                    help="run one call of the chosen workload after a 1 GiB device copy (for rocprofv3)")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("DIR", "NAME"),
                    help="write profiles/pmc_NAME.json from rocprofv3 counter CSVs under DIR")
    args = ap.parse_args()
    if args.summarize_pmc:
        return summarize_pmc(*args.summarize_pmc)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: the library has no multi-GPU path (asked for --gpus "
              f"{args.gpus}); the path would shard by z-slabs / query ranges with no exchange between GPUs",
              file=sys.stderr)
        sys.exit(2)
    if args.profile_run:
        return profile_run(args)
    args.sub_steps = max(1, min(args.steps, 3))

    import numpy as np
    import torch

    import meshlib_amd as M
    from meshlib_amd import synth

    torch.cuda.set_device(0)
    t0 = time.time()
    v, f = synth.bumpy_icosphere(MESH_LEVEL, MESH_BUMPS, MESH_SEED)
    dims, xf, vox = synth.grid_for_mesh(v, (GRID, GRID, GRID))
    pts = synth.projection_queries(v, f, NQ, QSEED)
    log(f"inputs {time.time() - t0:.1f}s: {f.shape[0]} tris, {GRID}^3 voxels, {NQ} queries")
    t0 = time.time()
    mesh = M.Mesh(v, f)
    depth = mesh.depth()
    log(f"tree build + upload {time.time() - t0:.1f}s (depth {depth})")
    fwn = M.FastWindingNumber(mesh)
    proj = M.PointsToMeshProjector()
    proj.updateMeshData(mesh)
    opts = M.DistanceToMeshOptions()
    nvox = GRID ** 3
    d_pts = torch.from_numpy(pts).cuda()
    d_sdf = torch.empty(nvox, dtype=torch.float32, device="cuda")
    d_proj = torch.empty(NQ * 8, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    timer = Timer(torch)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    sdf_ms, proj_ms = [], []

    def step():
        ev[0].record(stream)
        fwn.sdf_grid_device(dims, xf, opts, d_sdf.data_ptr(), stream=sh)
        ev[1].record(stream)
        proj.find_projections_device(d_pts.data_ptr(), NQ, d_proj.data_ptr(), stream=sh)
        ev[2].record(stream)

    M.set_traversal("wave_coop")
    for i in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log(f"warmup step {i} done")
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        step()
        torch.cuda.synchronize()
        sdf_ms.append(ev[0].elapsed_time(ev[1]))
        proj_ms.append(ev[1].elapsed_time(ev[2]))
        log(f"step {i}: sdf {sdf_ms[-1]:.2f} ms, projections {proj_ms[-1]:.2f} ms")
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t_start
    ms_per_step = 1000.0 * t_total / args.steps
    sdf_avg, proj_avg = float(np.mean(sdf_ms)), float(np.mean(proj_ms))
    res_sdf, res_proj = d_sdf.cpu(), d_proj.cpu()
    value = (nvox + NQ) / (ms_per_step / 1000.0)

    # the walks' own record counts for one call of each (results unchanged), for the rooflines
    cnt_sdf = counted(torch, M, [mesh], lambda: fwn.sdf_grid_device(dims, xf, opts, d_sdf.data_ptr(), stream=sh))
    cnt_proj = counted(torch, M, [mesh], lambda: proj.find_projections_device(d_pts.data_ptr(), NQ, d_proj.data_ptr(),
                                                                              stream=sh))
    same_counted = bool(torch.equal(d_sdf.cpu().view(torch.int32), res_sdf.view(torch.int32))
                        and torch.equal(d_proj.cpu(), res_proj))
    cb_sdf = with_bytes(cnt_sdf, io_bytes=4 * nvox)
    cb_proj = with_bytes(cnt_proj, io_bytes=48 * NQ)

    plain = None
    if not args.no_plain:
        M.set_traversal("per_lane")
        try:
            step()
            torch.cuda.synchronize()
        finally:
            M.set_traversal("wave_coop")
        p_sdf, p_proj = ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2])
        same_sdf = bool(torch.equal(d_sdf.cpu().view(torch.int32), res_sdf.view(torch.int32)))
        same_proj = bool(torch.equal(d_proj.cpu(), res_proj))
        plain = {"traversal": "per_lane", "sdf_ms": p_sdf, "proj_ms": p_proj,
                 "sdf_speedup": p_sdf / sdf_avg, "proj_speedup": p_proj / proj_avg,
                 "step_speedup": (p_sdf + p_proj) / (sdf_avg + proj_avg),
                 "identical_sdf": same_sdf, "identical_projections": same_proj}
        log(f"per-lane: sdf {p_sdf:.2f} ms, projections {p_proj:.2f} ms, identical {same_sdf} {same_proj}")

    # the host-pointer entry points (the reference's own signatures: inputs and outputs in host memory,
    # so the PCIe copies are inside the time); reported beside `value`, never as it
    host_ptr = {}
    for name, fn, ref in (("sdf", lambda: fwn.calcFromGridWithDistances(dims, xf, opts), res_sdf),
                          ("proj", lambda: proj.findProjections(pts), res_proj)):
        best = None
        for _ in range(2):
            t = time.perf_counter()
            r = fn()
            dt = 1e3 * (time.perf_counter() - t)
            best = dt if best is None else min(best, dt)
        host_ptr[name + "_ms"] = best
        host_ptr["identical_" + name] = bool(np.array_equal(r.view(np.uint8), ref.numpy().view(np.uint8)))
        del r
    host_ptr["queries_per_s"] = (nvox + NQ) / ((host_ptr["sdf_ms"] + host_ptr["proj_ms"]) / 1e3)
    host_ptr["note"] = ("calcFromGridWithDistances (67 MB out) and findProjections (50 MB in, 134 MB out) on "
                        "host arrays, best of 2, PCIe copies overlapped with the walks part by part")
    log(f"host-pointer: sdf {host_ptr['sdf_ms']:.1f} ms, projections {host_ptr['proj_ms']:.1f} ms, identical "
        f"{host_ptr['identical_sdf']} {host_ptr['identical_proj']}")

    ctx = {"M": M, "torch": torch, "np": np, "synth": synth, "v": v, "f": f, "mesh": mesh, "fwn": fwn,
           "proj": proj, "dims": dims, "xf": xf, "opts": opts, "d_pts": d_pts, "d_sdf": d_sdf, "d_proj": d_proj,
           "res_sdf": res_sdf, "res_proj": res_proj, "timer": timer, "sh": sh, "proj_avg": proj_avg,
           "sdf_avg": sdf_avg, "cnt_sdf": cnt_sdf, "cnt_proj": cnt_proj, "train_steps": args.train_steps}
    learned = None if args.no_learned else learned_leg(ctx)

    prof, pstate = pmc_profile("k8_step")
    roof_sdf = hbm_roofline("kSdfGridCoop+kSdfSplit+kSdfSplitFinish", sum(c["bytes"] for c in cb_sdf), sdf_avg,
                            cb_sdf, prof, pstate, ("kSdfGridCoop", "kSdfSplit"))
    roof_proj = hbm_roofline("kProjectionsCoop+kKeyedSplit+kProjSplitFinish", sum(c["bytes"] for c in cb_proj),
                             proj_avg, cb_proj, prof, pstate, ("kProjectionsCoop", "kKeyedSplit", "kProjSplitFinish"))
    rooflines = {"sdf": roof_sdf, "projections": roof_proj}
    if "mlp" in ctx:
        mlp = ctx["mlp"]
        for k, fl, ms in (("mlp_train", mlp["train_flop"], mlp["train_ms"]), ("mlp_infer", mlp["infer_flop"], mlp["infer_ms"])):
            ach = fl / (ms / 1e3) / 1e12
            rooflines[k] = {"bound": "mfma", "kernel": "BF16 MFMA MLP (mi_mlp.hip)", "achieved": ach,
                            "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": ach / BF16_PEAK_TFLOPS,
                            "traffic": None, "flop": fl, "ms": ms}

    batched = None if args.no_batched else batched_leg(args, M, torch, np, synth, timer)
    if batched:
        rooflines["batched"] = batched["roofline"]
    offset = None if args.no_offset else offset_leg(args, M, torch, np, synth, timer)
    if offset:
        rooflines["offset"] = offset["roofline"]

    cpu = None
    if not args.no_cpu:
        t = time.time()
        cpu = cpu_baseline(np, synth, v, f, dims, xf, pts, nvox)
        log(f"cpu baseline ({time.time() - t:.0f} s): {cpu['voxels_per_s']:.0f} vox/s, {cpu['pts_per_s']:.0f} pts/s "
            f"on {cpu['cores']} threads ({cpu['cpu_model']})")
    roof = {k: roof_sdf[k] for k in ("bound", "kernel", "achieved", "peak", "unit", "frac", "traffic")}
    roof.update({k: roof_sdf[k] for k in ("l2_hit_rate", "hbm_gbs_measured", "stall_frac", "valu_issue_frac",
                                            "pmc_profile", "bytes_per_call", "ms") if k in roof_sdf})
    out = {
        "metric": "signed-dist voxels/s & projected pts/s; learned-bound vs plain GPU speedup",
        "value": value, "unit": "queries/s (voxels + projected points)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "bumpy icosphere k=8 (1,310,720 tris, 32 bumps): 256^3 signed-distance "
                               "grid + 2^22 point projections per step", "voxels": nvox,
                   "points": NQ, "beta": 2.0},
        "voxels_per_s": nvox / (sdf_avg / 1000.0), "pts_per_s": NQ / (proj_avg / 1000.0),
        "sdf_ms": sdf_avg, "proj_ms": proj_avg, "traversal": "wave_coop", "vs_per_lane": plain,
        "identical_with_counters": same_counted, "host_pointer": host_ptr, "learned_vs_plain": learned,
        "batched": batched, "offset": offset,
        "roofline": roof, "rooflines": rooflines, "cpu_baseline": cpu,
    }
    if cpu:
        out["vs_cpu"] = value / cpu["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
