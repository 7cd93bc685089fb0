"""Throughput of the distance-field path on one MI355X.

Workload (the production shape): bumpy icosphere split 8 times (1,310,720 triangles, 32 bumps),
one step = a 256^3 signed-distance volume (calcFromGridWithDistances, FWN beta 2) + findProjections
of 2^22 query points (half uniform in the bbox grown 20 %, half in a Gaussian shell of sigma =
2 % of the bbox diagonal).  Inputs are resident in HBM before the timed region.

Prints ONE JSON line: value = (voxels + projected points) per second for the whole step, with the
per-part rates, the roofline of the dominant kernel (signed-distance grid, HBM-bound pointer
chasing; algorithmic bytes per voxel from the reference-order traversal counts of the CPU oracle
on a seeded sample) and the CPU baseline (the oracle, a C restatement of the reference CPU path,
timed on the host cores on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MESH_LEVEL, MESH_BUMPS, MESH_SEED = 8, 32, 2
GRID = 256
NQ = 1 << 22
QSEED = 7


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# HBM traffic of one signed-distance call on the default workload (its three launches: kSdfGridCoop,
# kSdfSplit, kSdfSplitFinish), from rocprofv3 --pmc passes over
# `bench.py --steps 1 --warmup 0 --no-cpu --no-learned --no-plain --copy-calibration`
# (profiles/sdf_grid_coop_k8_pmc.md): FETCH_SIZE and WRITE_SIZE in KiB summed over the three
# dispatches, each scaled by bytes / counter reading of the 1 GiB device copy in the same run.
PMC_SDF_KIB = {"fetch": 41281300.0 + 9248390.0 + 2261.06, "write": 139609.0 + 22481.8 + 5016.75}
PMC_COPY_KIB = {"fetch": 524308.0, "write": 1048576.0}
PMC_SDF_L2_HIT = 651274420.0 / (651274420.0 + 448592343.0)    # TCC_HIT_sum / (HIT + MISS), three launches
# SQ_INSTS_VALU of the same three dispatches (wave instructions), against the VALU issue rate of one
# wave64 instruction per SIMD every 4 cycles (1,024 SIMDs at 2.4 GHz peak clock).  The walks are bound
# by the latency of their dependent scalar loads rather than by issue: cutting kSdfGridCoop's VALU
# instructions by 23 % shortened it by 2 %, a seventh wave per SIMD by 7 %.
PMC_SDF_VALU = 3.02901e10 + 5.90839e9 + 5.69418e7
VALU_ISSUE_PER_S = 1024 * 2.4e9 / 4
COPY_BYTES = 1 << 30


def calibration_copy(torch):
    a = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b


def pmc_traffic_bytes():
    """HBM bytes per signed-distance launch, measured (fetch and write scaled separately)."""
    kib = COPY_BYTES / 1024.0
    return (PMC_SDF_KIB["fetch"] * kib / PMC_COPY_KIB["fetch"] + PMC_SDF_KIB["write"] * kib / PMC_COPY_KIB["write"]) * 1024.0


TRAIN_N, TRAIN_BATCH, TRAIN_LR, TRAIN_SEED = 1 << 20, 8192, 1e-3, 11
MARGINS = (0.05, 0.1, 0.25)
BATCH_TRAIN_N, BATCH_TRAIN_STEPS = 1 << 18, 500
# FLOP per sample: forward 2*(99*128 + 3*128*128 + 128); backward 2 GEMMs per layer (dW, dH)
MLP_FWD_FLOP = 2 * (99 * 128 + 3 * 128 * 128 + 128)
MLP_BWD_FLOP = 2 * (99 * 128 + 3 * 128 * 128 + 128) + 2 * (3 * 128 * 128 + 128)


def learned_leg(M, torch, np, synth, v, f, mesh, fwn, proj, dims, xf, opts, d_pts, d_sdf, d_proj,
                res_sdf, res_proj, sdf_avg, proj_avg, plain, ev, stream, sh):
    """Learned bound: train the MLP on labels from the GPU projector, then run projections and the
    signed-distance grid with the capped search at each margin; outputs must be identical."""
    NQ = d_pts.shape[0]
    nvox = d_sdf.shape[0]
    lo, hi = v.min(0).astype(np.float64), v.max(0).astype(np.float64)
    net = M.DistanceMLP(0.5 * (lo + hi), float(0.5 * np.max(hi - lo) * 1.2), seed=TRAIN_SEED)
    tp = torch.from_numpy(synth.projection_queries(v, f, TRAIN_N, TRAIN_SEED)).cuda()
    trec = torch.empty(TRAIN_N * 8, dtype=torch.int32, device="cuda")
    proj.find_projections_device(tp.data_ptr(), TRAIN_N, trec.data_ptr(), stream=sh)
    tlab = trec.view(torch.float32).view(-1, 8)[:, 7].sqrt().contiguous()
    torch.cuda.synchronize()
    loss0 = net.loss(tp[:65536], tlab[:65536])
    ev[0].record(stream)
    net.train(tp, tlab, steps=STEPS_TRAIN[0], batch=TRAIN_BATCH, lr=TRAIN_LR, stream=sh)
    ev[1].record(stream)
    torch.cuda.synchronize()
    train_ms = ev[0].elapsed_time(ev[1])
    loss1 = net.loss(tp[:65536], tlab[:65536])
    # held-out error of the prediction on the measured query set
    ref_d = res_proj.view(torch.float32).view(-1, 8)[:, 7].double().sqrt().numpy()
    ev[0].record(stream)
    pred = net.predict(d_pts)
    ev[1].record(stream)
    torch.cuda.synchronize()
    infer_ms = ev[0].elapsed_time(ev[1])
    rel = np.abs(pred - ref_d) / np.maximum(ref_d, 1e-3)
    out = {"train": {"samples": TRAIN_N, "steps": STEPS_TRAIN[0], "batch": TRAIN_BATCH, "lr": TRAIN_LR,
                     "ms": train_ms, "loss_before": loss0, "loss_after": loss1,
                     "tflops": STEPS_TRAIN[0] * TRAIN_BATCH * (MLP_FWD_FLOP + MLP_BWD_FLOP) / (train_ms / 1e3) / 1e12},
           "inference": {"points": NQ, "ms": infer_ms,
                         "tflops": NQ * MLP_FWD_FLOP / (infer_ms / 1e3) / 1e12,
                         "median_rel_err": float(np.median(rel)), "p90_rel_err": float(np.quantile(rel, 0.9)),
                         "under_prediction_rate": float(np.mean(pred < ref_d))},
           "margins": {}}
    base_step = sdf_avg + proj_avg
    for m in MARGINS:
        ev[0].record(stream)
        rp = proj.find_projections_learned_device(net, m, d_pts.data_ptr(), NQ, d_proj.data_ptr(), stream=sh)
        ev[1].record(stream)
        torch.cuda.synchronize()
        pm = ev[0].elapsed_time(ev[1])
        same_p = bool(torch.equal(d_proj.cpu(), res_proj))
        ev[0].record(stream)
        rs = fwn.sdf_grid_learned_device(net, m, dims, xf, opts, d_sdf.data_ptr(), stream=sh)
        ev[1].record(stream)
        torch.cuda.synchronize()
        sm = ev[0].elapsed_time(ev[1])
        same_s = bool(torch.equal(d_sdf.cpu().view(torch.int32), res_sdf.view(torch.int32)))
        e = {"proj_ms": pm, "sdf_ms": sm, "proj_fallback_rate": rp / NQ, "sdf_fallback_rate": rs / nvox,
             "identical_projections": same_p, "identical_sdf": same_s,
             "speedup_vs_plain_coop": base_step / (pm + sm)}
        if plain:
            e["speedup_vs_per_lane"] = (plain["sdf_ms"] + plain["proj_ms"]) / (pm + sm)
        out["margins"][str(m)] = e
        log(f"learned m={m}: proj {pm:.2f} ms (fallback {rp / NQ:.4f}), sdf {sm:.2f} ms "
            f"(fallback {rs / nvox:.4f}), identical {same_p} {same_s}")
    return out


STEPS_TRAIN = [2000]


def batched_main(args):
    """Uneven batch: 64 bumpy icospheres with split level uniform in 4..8, each with a signed-distance
    grid of side in {32, 64, 96, 128, 192} and 2^18..2^21 projection queries.  Compares one
    autotuned persistent launch with per-item launches (wave-cooperative and per-lane forms)."""
    import numpy as np
    import torch

    import meshlib_amd as M
    from meshlib_amd import synth

    torch.cuda.set_device(0)
    rng = np.random.Generator(np.random.PCG64(4))
    n_items = args.items
    levels = rng.integers(4, 9, size=n_items)
    sides = rng.choice([32, 64, 96, 128, 192], size=n_items)
    nqs = rng.integers(1 << 18, (1 << 21) + 1, size=n_items)
    t0 = time.time()
    items, keep, per_item, geoms = [], [], [], []
    tot_vox = tot_pts = 0
    for i in range(n_items):
        v, f = synth.bumpy_icosphere(int(levels[i]), 32, 1000 + i)
        mesh = M.Mesh(v, f)
        dims, xf, vox = synth.grid_for_mesh(v, (int(sides[i]),) * 3)
        pts = synth.projection_queries(v, f, int(nqs[i]), 2000 + i)
        dp = torch.from_numpy(pts).cuda()
        ds = torch.empty(int(np.prod(dims)), dtype=torch.float32, device="cuda")
        dr = torch.empty(int(nqs[i]) * 8, dtype=torch.int32, device="cuda")
        items.append({"mesh": mesh, "dims": dims, "xf": xf, "d_sdf": ds.data_ptr(), "d_pts": dp.data_ptr(),
                      "n": int(nqs[i]), "d_proj": dr.data_ptr()})
        fwn = M.FastWindingNumber(mesh)
        pr = M.PointsToMeshProjector()
        pr.updateMeshData(mesh)
        per_item.append((fwn, pr, dims, xf, ds, dp, dr))
        geoms.append((v, f))
        keep += [dp, ds, dr]
        tot_vox += int(np.prod(dims))
        tot_pts += int(nqs[i])
    log(f"batched inputs {time.time() - t0:.1f}s: {n_items} meshes, {tot_vox} voxels, {tot_pts} points")
    opts = M.DistanceToMeshOptions()
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def timed(fn):
        ev[0].record(stream)
        fn()
        ev[1].record(stream)
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1])

    def per_item_run():
        for fwn, pr, dims, xf, ds, dp, dr in per_item:
            fwn.sdf_grid_device(dims, xf, opts, ds.data_ptr())
            pr.find_projections_device(dp.data_ptr(), dp.shape[0], dr.data_ptr())

    M.set_traversal("wave_coop")
    per_item_run()
    torch.cuda.synchronize()
    t_items = timed(per_item_run)
    ref = [(ds.clone(), dr.clone()) for (_, _, _, _, ds, _, dr) in per_item]
    batch = M.Batch()
    t_tune = batch.autotune(items, opts)
    cfg = batch.config()
    t_batch = min(timed(lambda: batch.run(items, opts)) for _ in range(args.steps))
    same = all(torch.equal(ds.view(torch.int32), r[0].view(torch.int32)) and torch.equal(dr, r[1])
               for (_, _, _, _, ds, _, dr), r in zip(per_item, ref))
    learned = None
    if not args.no_learned:
        # one network per mesh, trained on the GPU from its own projector's distances (2^18 samples,
        # 500 Adam steps each), then the batch with caps (pred * (1 + m))^2 and exact reruns
        t0 = time.time()
        nets = []
        for i, ((fwn, pr, dims, xf, ds, dp, dr), (v, f)) in enumerate(zip(per_item, geoms)):
            lo, hi = v.min(0).astype(np.float64), v.max(0).astype(np.float64)
            net = M.DistanceMLP(0.5 * (lo + hi), float(0.5 * np.max(hi - lo) * 1.2), seed=TRAIN_SEED + i)
            tp = torch.from_numpy(synth.projection_queries(v, f, BATCH_TRAIN_N, 3000 + i)).cuda()
            trec = torch.empty(BATCH_TRAIN_N * 8, dtype=torch.int32, device="cuda")
            pr.find_projections_device(tp.data_ptr(), BATCH_TRAIN_N, trec.data_ptr())
            tlab = trec.view(torch.float32).view(-1, 8)[:, 7].sqrt().contiguous()
            net.train(tp, tlab, steps=BATCH_TRAIN_STEPS, batch=TRAIN_BATCH, lr=TRAIN_LR)
            nets.append(net)
        torch.cuda.synchronize()
        train_s = time.time() - t0
        learned = {"train_s_all_meshes": train_s, "samples_per_mesh": BATCH_TRAIN_N, "steps_per_mesh": BATCH_TRAIN_STEPS,
                   "margins": {}}
        for m in (0.1, 0.25):
            reruns = [0]

            def go():
                reruns[0] = batch.run_learned(items, nets, m, opts)
            t_l = min(timed(go) for _ in range(args.steps))
            same_l = all(torch.equal(ds.view(torch.int32), r[0].view(torch.int32)) and torch.equal(dr, r[1])
                         for (_, _, _, _, ds, _, dr), r in zip(per_item, ref))
            learned["margins"][str(m)] = {"ms": t_l, "reruns": reruns[0],
                                          "rerun_rate": reruns[0] / (tot_vox + tot_pts),
                                          "identical_to_per_item": bool(same_l),
                                          "speedup_vs_batched_plain": t_batch / t_l,
                                          "speedup_vs_per_item_coop": t_items / t_l}
            log(f"learned+batched m={m}: {t_l:.1f} ms, reruns {reruns[0]}, identical {same_l}")
    t_plain = None
    if not args.no_plain:
        M.set_traversal("per_lane")
        t_plain = timed(per_item_run)
        M.set_traversal("wave_coop")
        if learned:
            for e in learned["margins"].values():
                e["speedup_vs_per_item_per_lane"] = t_plain / e["ms"]
    out = {"metric": "signed-dist voxels/s & projected pts/s (batched uneven workload)",
           "value": (tot_vox + tot_pts) / (t_batch / 1e3), "unit": "queries/s (voxels + projected points)",
           "n_gpus": 1, "steps": args.steps, "ms_per_step": t_batch, "higher_is_better": True,
           "data": "synthetic", "dtype": "f32+f64",
           "config": {"workload": f"{n_items} bumpy icospheres, levels 4..8, grids 32..192, 2^18..2^21 points each",
                      "voxels": tot_vox, "points": tot_pts},
           "batched_ms": t_batch, "autotune": {"blocks_per_cu": cfg[0], "xcd_split": cfg[1], "best_ms": t_tune},
           "per_item_coop_ms": t_items, "per_item_per_lane_ms": t_plain,
           "speedup_vs_per_item_coop": t_items / t_batch,
           "speedup_vs_per_item_per_lane": (t_plain / t_batch) if t_plain else None,
           "identical_to_per_item": bool(same), "learned_batched": learned}
    print(json.dumps(out), flush=True)


def offset_main(args):
    """Offset-scale volume: bumpy icosphere split 9 times (5,242,880 tris, 64 bumps), 512^3 signed
    distances (FWN beta 2); plain wave-cooperative traversal against the learned-bound form at each
    margin (outputs identical), and one run of the per-lane form unless --no-plain."""
    import numpy as np
    import torch

    import meshlib_amd as M
    from meshlib_amd import synth

    torch.cuda.set_device(0)
    t0 = time.time()
    v, f = synth.bumpy_icosphere(9, 64, 5)
    dims, xf, vox = synth.grid_for_mesh(v, (512, 512, 512))
    nvox = 512 ** 3
    log(f"inputs {time.time() - t0:.1f}s: {f.shape[0]} tris, 512^3 voxels")
    t0 = time.time()
    mesh = M.Mesh(v, f)
    log(f"tree build + upload {time.time() - t0:.1f}s (depth {mesh.depth()})")
    fwn = M.FastWindingNumber(mesh)
    proj = M.PointsToMeshProjector()
    proj.updateMeshData(mesh)
    opts = M.DistanceToMeshOptions()
    d_sdf = torch.empty(nvox, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def timed(fn):
        ev[0].record(stream)
        r = fn()
        ev[1].record(stream)
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1]), r

    M.set_traversal("wave_coop")
    for _ in range(args.warmup):
        timed(lambda: fwn.sdf_grid_device(dims, xf, opts, d_sdf.data_ptr(), stream=sh))
    times = []
    for i in range(args.steps):
        t, _ = timed(lambda: fwn.sdf_grid_device(dims, xf, opts, d_sdf.data_ptr(), stream=sh))
        times.append(t)
        log(f"step {i}: sdf {t:.1f} ms")
    sdf_ms = float(np.mean(times))
    ref = d_sdf.view(torch.int32).clone()
    plain = None
    if not args.no_plain:
        M.set_traversal("per_lane")
        t, _ = timed(lambda: fwn.sdf_grid_device(dims, xf, opts, d_sdf.data_ptr(), stream=sh))
        M.set_traversal("wave_coop")
        plain = {"sdf_ms": t, "speedup": t / sdf_ms, "identical": bool(torch.equal(d_sdf.view(torch.int32), ref))}
        log(f"per-lane: {t:.1f} ms")

    lo, hi = v.min(0).astype(np.float64), v.max(0).astype(np.float64)
    net = M.DistanceMLP(0.5 * (lo + hi), float(0.5 * np.max(hi - lo) * 1.2), seed=TRAIN_SEED)
    tp = torch.from_numpy(synth.projection_queries(v, f, TRAIN_N, TRAIN_SEED)).cuda()
    trec = torch.empty(TRAIN_N * 8, dtype=torch.int32, device="cuda")
    proj.find_projections_device(tp.data_ptr(), TRAIN_N, trec.data_ptr(), stream=sh)
    tlab = trec.view(torch.float32).view(-1, 8)[:, 7].sqrt().contiguous()
    train_ms, _ = timed(lambda: net.train(tp, tlab, steps=STEPS_TRAIN[0], batch=TRAIN_BATCH, lr=TRAIN_LR,
                                          stream=sh))
    learned = {}
    for m in MARGINS:
        t, r = timed(lambda: fwn.sdf_grid_learned_device(net, m, dims, xf, opts, d_sdf.data_ptr(), stream=sh))
        same = bool(torch.equal(d_sdf.view(torch.int32), ref))
        learned[str(m)] = {"sdf_ms": t, "fallback_rate": r / nvox, "identical": same,
                           "speedup_vs_plain_coop": sdf_ms / t}
        log(f"learned m={m}: {t:.1f} ms, fallback {r / nvox:.4f}, identical {same}")
    out = {"metric": "signed-dist voxels/s (offset-scale volume)", "value": nvox / (sdf_ms / 1e3),
           "unit": "voxels/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": sdf_ms,
           "higher_is_better": True, "data": "synthetic", "dtype": "f32+f64",
           "config": {"workload": "bumpy icosphere k=9 (5,242,880 tris, 64 bumps): 512^3 signed-distance grid",
                      "voxels": nvox, "beta": 2.0},
           "vs_per_lane": plain,
           "learned_vs_plain": {"train_ms": train_ms, "train_steps": STEPS_TRAIN[0], "margins": learned}}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--traversal", default="wave_coop", choices=["wave_coop", "per_lane"],
                    help="kernel traversal form of the measured (value) run")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the one-step comparison run with the per-lane (reference-form) traversal")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-bound leg")
    ap.add_argument("--train-steps", type=int, default=2000)
    ap.add_argument("--workload", default="single", choices=["single", "batched", "offset"],
                    help="single: the k=8 mesh step (default); batched: the uneven multi-mesh batch; "
                         "offset: the k=9 512^3 volume")
    ap.add_argument("--items", type=int, default=64, help="meshes in the batched workload")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 1 GiB device-to-device copy first (reference kernel for PMC byte counters)")
    args = ap.parse_args()
    STEPS_TRAIN[0] = args.train_steps
    if args.workload == "batched" and args.gpus == 1:
        return batched_main(args)
    if args.workload == "offset" and args.gpus == 1:
        return offset_main(args)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (asked for --gpus "
              f"{args.gpus}); the path shards by z-slabs / query ranges with no exchange, see DESIGN.md",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import meshlib_amd as M
    from meshlib_amd import synth

    torch.cuda.set_device(0)
    if args.copy_calibration:
        calibration_copy(torch)
    t0 = time.time()
    v, f = synth.bumpy_icosphere(MESH_LEVEL, MESH_BUMPS, MESH_SEED)
    dims, xf, vox = synth.grid_for_mesh(v, (GRID, GRID, GRID))
    pts = synth.projection_queries(v, f, NQ, QSEED)
    log(f"inputs {time.time() - t0:.1f}s: {f.shape[0]} tris, {GRID}^3 voxels, {NQ} queries")
    t0 = time.time()
    mesh = M.Mesh(v, f)
    log(f"tree build + upload {time.time() - t0:.1f}s (depth {mesh.depth()})")
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

    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    sdf_ms, proj_ms = [], []

    def step(timed):
        ev[0].record(stream)
        fwn.sdf_grid_device(dims, xf, opts, d_sdf.data_ptr(), stream=sh)
        ev[1].record(stream)
        proj.find_projections_device(d_pts.data_ptr(), NQ, d_proj.data_ptr(), stream=sh)
        ev[2].record(stream)
        if timed:
            return True

    M.set_traversal(args.traversal)
    for i in range(args.warmup):
        step(False)
        torch.cuda.synchronize()
        log(f"warmup step {i} done")
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        step(True)
        torch.cuda.synchronize()
        sdf_ms.append(ev[0].elapsed_time(ev[1]))
        proj_ms.append(ev[1].elapsed_time(ev[2]))
        log(f"step {i}: sdf {sdf_ms[-1]:.2f} ms, projections {proj_ms[-1]:.2f} ms")
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t_start
    ms_per_step = 1000.0 * t_total / args.steps
    sdf_avg = float(np.mean(sdf_ms))
    proj_avg = float(np.mean(proj_ms))
    res_sdf = d_sdf.cpu()
    res_proj = d_proj.cpu()

    # one step with the per-lane traversal (the reference's kernel form) for the speedup figure;
    # outputs must be identical
    plain = None
    if not args.no_plain and args.traversal != "per_lane":
        M.set_traversal("per_lane")
        step(True)
        torch.cuda.synchronize()
        p_sdf, p_proj = ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2])
        same_sdf = bool(torch.equal(d_sdf.cpu().view(torch.int32), res_sdf.view(torch.int32)))
        same_proj = bool(torch.equal(d_proj.cpu(), res_proj))
        M.set_traversal(args.traversal)
        plain = {"traversal": "per_lane", "sdf_ms": p_sdf, "proj_ms": p_proj,
                 "sdf_speedup": p_sdf / sdf_avg, "proj_speedup": p_proj / proj_avg,
                 "step_speedup": (p_sdf + p_proj) / (sdf_avg + proj_avg),
                 "identical_sdf": same_sdf, "identical_projections": same_proj}
        log(f"per-lane: sdf {p_sdf:.2f} ms, projections {p_proj:.2f} ms, identical {same_sdf} {same_proj}")
    vox_rate = nvox / (sdf_avg / 1000.0)
    pts_rate = NQ / (proj_avg / 1000.0)
    value = (nvox + NQ) / (ms_per_step / 1000.0)

    learned = None
    if not args.no_learned:
        learned = learned_leg(M, torch, np, synth, v, f, mesh, fwn, proj, dims, xf, opts, d_pts, d_sdf, d_proj,
                              res_sdf, res_proj, sdf_avg, proj_avg, plain, ev, stream, sh)

    # ---- CPU baseline (the oracle: C restatement of the reference CPU path) + algorithmic bytes
    cpu = None
    b_vox = b_pt = None
    if not args.no_cpu:
        from oracle import oracle as O
        cores = max(1, min(16, os.cpu_count() or 1))
        om = O.Mesh(v, f)
        rng = np.random.default_rng(123)
        budget = args.cpu_seconds / 2
        cv, cp = O.Counters(), O.Counters()
        n_v = n_p = 0
        t_v = t_p = 0.0
        while t_v < budget:
            ids = rng.choice(nvox, 2000 * cores, replace=False).astype(np.int64)
            t = time.perf_counter()
            om.sdf_grid_sample(dims, xf, ids, nthreads=cores, counters=cv)
            t_v += time.perf_counter() - t
            n_v += ids.size
        while t_p < budget:
            sel = rng.choice(NQ, 4000 * cores, replace=False)
            t = time.perf_counter()
            om.find_projections(pts[sel], nthreads=cores, counters=cp)
            t_p += time.perf_counter() - t
            n_p += sel.size
        cpu_vox_rate = n_v / t_v
        cpu_pts_rate = n_p / t_p
        # the same sample kind on one thread, for the per-core figure
        ids1 = rng.choice(nvox, 4000, replace=False).astype(np.int64)
        t = time.perf_counter()
        om.sdf_grid_sample(dims, xf, ids1, nthreads=1)
        vox1 = ids1.size / (time.perf_counter() - t)
        sel1 = rng.choice(NQ, 8000, replace=False)
        t = time.perf_counter()
        om.find_projections(pts[sel1], nthreads=1)
        pts1 = sel1.size / (time.perf_counter() - t)
        cpu_step = nvox / cpu_vox_rate + NQ / cpu_pts_rate
        cv_d, cp_d = cv.as_dict(), cp.as_dict()
        # algorithmic bytes per unit (reference AoS record sizes): query 12 + result 32, 32 B per
        # child-box probe, 48 B per leaf test; voxel: 4 B out, the distance part, and 64 B per FWN
        # node visit + 48 B per solid-angle leaf
        b_pt = 12 + 32 + (32 * cp_d["dist_probes"] + 48 * cp_d["dist_leaves"]) / cp_d["queries"]
        b_vox = 4 + (32 * cv_d["dist_probes"] + 48 * cv_d["dist_leaves"] + 64 * cv_d["fwn_visits"]
                     + 48 * cv_d["fwn_leaves"]) / cv_d["queries"]
        cpu = {"value": (nvox + NQ) / cpu_step, "unit": "queries/s", "cores": cores, "kind": "port",
               "sample": f"{n_v} seeded voxels of the 256^3 grid + {n_p} seeded queries of the 2^22 set, "
                         f"{t_v + t_p:.1f}s of CPU time; step time extrapolated as voxels/rate + points/rate",
               "voxels_per_s": cpu_vox_rate, "pts_per_s": cpu_pts_rate,
               "one_thread": {"voxels_per_s": vox1, "pts_per_s": pts1,
                              "sample": f"{ids1.size} voxels + {sel1.size} queries"}}
        log(f"cpu baseline: {cpu_vox_rate:.0f} vox/s, {cpu_pts_rate:.0f} pts/s on {cores} threads")

    roof = None
    if b_vox is not None:
        achieved = b_vox * nvox / (sdf_avg / 1000.0) / 1e9
        kname = ("kSdfGridCoop+kSdfSplit+kSdfSplitFinish" if args.traversal == "wave_coop" else "kSdfGrid")
        traffic = pmc_traffic_bytes() if args.traversal == "wave_coop" else None
        roof = {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "bytes_per_voxel": b_vox, "bytes_per_point": b_pt,
                "algorithmic_bytes_per_launch": b_vox * nvox,
                "hbm_gbs_measured": (traffic / (sdf_avg / 1000.0) / 1e9) if traffic else None,
                "l2_hit_rate": PMC_SDF_L2_HIT if traffic else None,
                "valu_insts_per_launch": PMC_SDF_VALU if traffic else None,
                "valu_issue_frac": (PMC_SDF_VALU / (sdf_avg / 1000.0) / VALU_ISSUE_PER_S) if traffic else None}
    out = {
        "metric": "signed-dist voxels/s & projected pts/s; learned-bound vs plain GPU speedup",
        "value": value, "unit": "queries/s (voxels + projected points)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "bumpy icosphere k=8 (1,310,720 tris, 32 bumps): 256^3 signed-distance "
                               "grid + 2^22 point projections per step", "voxels": nvox,
                   "points": NQ, "beta": 2.0},
        "voxels_per_s": vox_rate, "pts_per_s": pts_rate, "sdf_ms": sdf_avg, "proj_ms": proj_avg,
        "traversal": args.traversal, "vs_per_lane": plain,
        "learned_vs_plain": learned,
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
