"""Throughput benchmark: primary rays per second on the 512^3 displaced-icosphere SBS.

Workload (one step): 8 seeded orbit views x 2048x2048 primary rays
(33,554,432 rays) through the BVH2 over ~1e5 bricks of the displaced
icosphere (4 noise octaves, 512^3 effective = brick LOD 128 x 4^3 voxels,
8-bit distances over the voxel diagonal), Newton root find, depth + normal
G-buffer written to HBM. Scene build and upload happen before the timed
region; each step is one kernel launch.

Prints one JSON line. `roofline` uses the algorithmic bytes per launch from
the kernel's own work counters: 64 B per BVH node pair, 16 B per brick record,
8 x bytes_per_value per voxel-corner fetch and 16 B of output per ray, divided
by the kernel's average duration (HIP events on the launch stream).
`cpu_baseline` times the CPU oracle (the reference CPU path restated in C++,
FP32, OpenMP over the host cores) on the same workload. The DR leg
(`dr_fwd_bwd_mrays_s`) times RenderDR + Adam on the torus-from-sphere problem (torus target,
sphere start, 128^3 indexed SBS, 16 views at 512x512, border_spp 256, 200
Adam steps at lr 1e-3). The `phong_shadow` leg renders the plain views with
Phong shading, one shadow ray toward the direct light per hit.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

W = H = 2048
N_VIEWS = 8
SCENE_SEED = 2
CAM_SEED = 2
BATCH_SEED = 5
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def measured_counters(lod):
    """The plain-render kernel's counter figures recorded per scene size in
    profiles/k_render_ico<4*lod>_hbm.json by --reduce-counters: HBM bytes per
    launch (FETCH_SIZE + WRITE_SIZE scaled by a 1 GiB copy), VALU lane
    utilisation and VALU issue fraction ({} when no pass was recorded)."""
    path = os.path.join(ROOT, "profiles", f"k_render_ico{4 * lod}_hbm.json")
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def reduce_counters(dirs, kernel, rays, calib_bytes, out_path):
    """Reduces rocprofv3 --pmc passes (csv output, one pass per directory or
    several in one tree) to the per-launch figures bench.py reports: the
    average over the dispatches of `kernel` of every counter, FETCH_SIZE /
    WRITE_SIZE scaled to bytes by what the same counters read for the HIP
    runtime's device-to-device copy of `calib_bytes` (bench.py --calib-copy-mib),
    the L2 hit rate, VALU lane utilisation, VALU instructions per 64-ray wave
    and resident waves per CU. Writes the JSON bench.py reads as `traffic`."""
    import csv
    import glob
    per = {}    # counter -> list of per-dispatch values of the kernel
    calib = {}  # counter -> list of per-dispatch values of the copy kernel
    names = set()
    for d in dirs:
        for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            acc = {}
            with open(path, newline="") as f:
                for row in csv.DictReader(f):
                    name = row.get("Kernel_Name", "")
                    key = (name, row.get("Dispatch_Id", ""), row["Counter_Name"])
                    acc[key] = acc.get(key, 0.0) + float(row["Counter_Value"])
            for (name, _, counter), v in acc.items():
                if kernel in name:
                    per.setdefault(counter, []).append(v)
                    names.add(name)
                elif "copyBuffer" in name:
                    calib.setdefault(counter, []).append(v)
    if not per:
        raise SystemExit(f"no dispatch of a kernel matching {kernel!r} in {dirs}")
    mean = {k: sum(v) / len(v) for k, v in per.items()}
    out = {"kernel": sorted(names), "dispatches": {k: len(v) for k, v in per.items()}, "counters_per_dispatch": mean}
    if "FETCH_SIZE" in mean and "WRITE_SIZE" in mean and "FETCH_SIZE" in calib and "WRITE_SIZE" in calib:
        fs = calib_bytes / 1024.0 / max(calib["FETCH_SIZE"])   # KiB read by the copy vs KiB counted
        ws = calib_bytes / 1024.0 / max(calib["WRITE_SIZE"])
        rd, wr = mean["FETCH_SIZE"] * fs * 1024, mean["WRITE_SIZE"] * ws * 1024
        out.update({"calibration": {"bytes": calib_bytes, "fetch_size_kib": max(calib["FETCH_SIZE"]),
                                    "write_size_kib": max(calib["WRITE_SIZE"]), "fetch_scale": round(fs, 4),
                                    "write_scale": round(ws, 4)},
                    "hbm_read_bytes": int(rd), "hbm_write_bytes": int(wr), "traffic_bytes": int(rd + wr)})
    if "TCC_HIT_sum" in mean and "TCC_MISS_sum" in mean:
        out["l2_hit_rate"] = round(mean["TCC_HIT_sum"] / (mean["TCC_HIT_sum"] + mean["TCC_MISS_sum"]), 4)
    if "SQ_THREAD_CYCLES_VALU" in mean and "SQ_ACTIVE_INST_VALU" in mean:
        out["valu_lane_utilisation"] = round(mean["SQ_THREAD_CYCLES_VALU"] / (64.0 * mean["SQ_ACTIVE_INST_VALU"]), 4)
    if "SQ_INSTS_VALU" in mean and rays:
        out["valu_insts_per_wave_tile"] = round(mean["SQ_INSTS_VALU"] / (rays / 64.0), 1)
    if "GRBM_GUI_ACTIVE" in mean:
        # GRBM_GUI_ACTIVE sums the busy cycles of the 8 XCDs; the SQ cycle
        # counters count in units of 4 cycles, summed over waves (SQ_WAVE_CYCLES)
        # or over the SIMDs' VALU issue (SQ_ACTIVE_INST_VALU)
        cycles = mean["GRBM_GUI_ACTIVE"] / 8.0
        out["kernel_cycles"] = round(cycles)
        if "SQ_WAVE_CYCLES" in mean:
            out["waves_per_cu"] = round(4.0 * mean["SQ_WAVE_CYCLES"] / cycles / 256.0, 2)
        if "SQ_ACTIVE_INST_VALU" in mean:
            out["valu_issue_frac"] = round(4.0 * mean["SQ_ACTIVE_INST_VALU"] / (cycles * 1024.0), 4)
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


def plain_scene(L, lod=128):
    """The displaced icosphere (4 noise octaves, 8-bit distances over the voxel
    diagonal) at brick LOD `lod` x 4^3 voxels; 5 subdivisions up to LOD 128,
    6 above so the mesh stays finer than the voxels."""
    return L.displaced_icosphere(5 if lod <= 128 else 6, 0.8, 0.05, 4, seed=SCENE_SEED, lod=lod, brick_size=4, brick_pad=0,
                                 bytes_per_value=1, voxel_diag_scale=True)


def dr_problem(L, np, views=16, res=512, brick_count=32, brick_size=4):
    """DR problem: noisy torus (R 0.6, r 0.2) target, sphere r 0.5 start, dense
    indexed SBS of 128^3 voxels, `views` seeded views at res x res."""
    from litert_amd import workloads as WL
    tgt, ini = WL.dr_scenes(brick_count, brick_size)
    cams = WL.dr_cameras(views)
    r = L.CreateMultiRenderer(L.DEVICE_GPU)
    r.SetScene(tgt)
    p = WL.dr_reference_preset()
    refs = np.stack([r.RenderFloat(res, res, wv, pr, p) for wv, pr in cams])
    r.close()
    return tgt, ini, cams, refs, WL.dr_preset()


def time_views(r, wvs, prs, steps, warmup):
    """Wall-clock Mrays/s and mean kernel ms of `steps` launches of all views."""
    for _ in range(warmup):
        r.RenderViewsDevice(wvs, prs, W, H, 0, timed=True)
    r.Sync()
    ks = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ks.append(r.RenderViewsDevice(wvs, prs, W, H, 0, timed=True))
    r.Sync()
    el = time.perf_counter() - t0
    return wvs.shape[0] * W * H * steps / el / 1e6, sum(ks) / len(ks)


def mlp_leg(L, np, r, wvs, prs, args):
    """Learned seeder: collect (ray, t_hit) events from plain renders of 32
    other seeded 128^3 FP32 displaced icospheres, train on the GPU, then time
    seeded vs plain on (a) the 128^3 FP32 icosphere with the reference's default
    sphere-tracing root finder and (b) the 512^3 8-bit icosphere (Newton)."""
    MLP_FLOP = 19840
    t0 = time.perf_counter()
    r.MlpDatasetClear()
    per_scene = args.mlp_events // args.mlp_scenes
    for k in range(args.mlp_scenes):
        s = L.displaced_icosphere(5, 0.8, 0.05, 3, seed=1000 + k, lod=32, brick_size=4, brick_pad=1, bytes_per_value=4)
        r.SetScene(s)
        cams = [L.orbit_camera(5000 + k, i) for i in range(2)]
        r.MlpCollect(np.stack([c[0] for c in cams]), np.stack([c[1] for c in cams]), 512, 512, per_scene)
    n = r.MlpDatasetSize()
    collect_s = time.perf_counter() - t0
    r.MlpInit(1)
    losses, train_ms = r.MlpTrain(args.mlp_epochs, batch=16384, lr=2e-3, seed=1)
    f, t, _ = r.MlpDataset(0, min(n, 200000))
    out, _ = r.MlpInfer(f)
    acc = float(np.mean((out[:, 2] > 0) == (t[:, 1] > 0.5)))
    hit = t[:, 1] > 0.5
    t_err = float(np.median(np.abs(0.5 * (out[hit, 0] + out[hit, 1]) - t[hit, 0]))) if hit.any() else None
    inf_ms, _ = r.MlpInferDataset(repeats=5)
    tflops = MLP_FLOP * n / (inf_ms * 1e-3) / 1e12
    log(f"mlp: {n} events from {args.mlp_scenes} scenes ({collect_s:.1f}s), train {train_ms / 1e3:.2f}s, loss "
        f"{losses[0]:.4f}->{losses[-1]:.4f}, hit acc {acc:.4f}, inference {inf_ms:.3f} ms ({tflops:.1f} TFLOP/s)")
    info = {"train_scenes": args.mlp_scenes, "train_events": int(n), "epochs": args.mlp_epochs, "train_ms": round(train_ms, 1),
            "train_samples_per_s": round(n * args.mlp_epochs / (train_ms * 1e-3), 0),
            "loss_first_last": [float(losses[0]), float(losses[-1])], "hit_accuracy": round(acc, 4),
            "median_abs_t_err_brick_units": t_err,
            "inference": {"samples": int(n), "kernel_ms": round(inf_ms, 4), "tflops": round(tflops, 2),
                          "mfma_bf16_peak_tflops": 2500.0, "mfma_frac": round(tflops / 2500.0, 5)}}
    res = {}
    for name, scene, mode in (("ico128_fp32_st", "ico128", L.SDF_OCTREE_NODE_INTERSECT_ST),
                              ("ico512_u8_newton", "ico512", L.SDF_OCTREE_NODE_INTERSECT_NEWTON)):
        sbs = (L.displaced_icosphere(5, 0.8, 0.05, 3, seed=1, lod=32, brick_size=4, brick_pad=1, bytes_per_value=4)
               if scene == "ico128" else plain_scene(L))
        r.SetScene(sbs)
        pp = L.getDefaultPreset()
        pp.sdf_node_intersect = mode
        pp.render_mode = L.MULTI_RENDER_MODE_NORMAL
        r.SetPreset(pp)
        row = {}
        for seeded in (False, True):
            r.SetSeeding(seeded, 0.9)
            mr, kms = time_views(r, wvs, prs, args.steps, 1)
            cnt = r.CountWorkEx(wvs, prs, W, H)
            key = "seeded" if seeded else "plain"
            row[key + "_mrays_s"] = round(mr, 1)
            row[key + "_kernel_ms"] = round(kms, 3)
            row[key + "_root_iters_per_ray"] = round(float(cnt[3]) / (wvs.shape[0] * W * H), 4)
            if seeded:
                row["mlp_tiles"] = int(cnt[4])
                row["gated_bricks"] = int(cnt[5])
        r.SetSeeding(False)
        row["bricks"] = int(sbs.nodes.shape[0])
        res[name] = row
        log(f"{name}: plain {row['plain_mrays_s']} Mrays/s, seeded {row['seeded_mrays_s']} Mrays/s; root iters/ray "
            f"{row['plain_root_iters_per_ray']} -> {row['seeded_root_iters_per_ray']}")
    info.update(res)
    return res["ico512_u8_newton"]["seeded_mrays_s"], info


def batch_leg(L, np, args, mlp_params):
    """Uneven batch: `args.batch_scenes` seeded scenes (displaced icosphere or
    torus; octree depth d ~ U{5..9}, i.e. brick LOD 2^(d-2) x 4^3 voxels; 8-bit
    distances over the voxel diagonal or FP32), each with its own seeded orbit
    camera and image of W x H ~ U[128, 2048]^2, rendered in one persistent
    launch (Newton, depth + normal) with the chunk and grid sizes picked by the
    autotuner; compared with the same items rendered one launch per scene
    (the same batched kernel given one item at a time, so the plain
    single-scene kernel's statistics cover the 512^3 workload only). With a
    trained network the seeded form of the same batch is timed too."""
    from litert_amd import workloads as WL
    p = L.getDefaultPreset()
    p.sdf_node_intersect = L.SDF_OCTREE_NODE_INTERSECT_NEWTON
    p.render_mode = L.MULTI_RENDER_MODE_NORMAL
    rs, wvs, prs, ws, hs, bricks = [], [], [], [], [], []
    t0 = time.perf_counter()
    for spec in WL.batch_specs(args.batch_scenes, BATCH_SEED):
        s = WL.batch_scene(spec)
        r = L.CreateMultiRenderer(L.DEVICE_GPU)
        r.SetScene(s)
        r.SetPreset(p)
        bricks.append(int(s.nodes.shape[0]))
        del s
        wv, pr = WL.batch_camera(spec)
        rs.append(r), wvs.append(wv), prs.append(pr), ws.append(spec["w"]), hs.append(spec["h"])
    build_s = time.perf_counter() - t0
    wvs, prs = np.stack(wvs), np.stack(prs)
    rays = int(sum(w * h for w, h in zip(ws, hs)))
    owner = L.CreateMultiRenderer(L.DEVICE_GPU)
    owner.SetPreset(p)

    def timed_batch():
        owner.RenderBatch(rs, wvs, prs, ws, hs, autotune=True)
        chunk, blocks = owner.BatchConfig()
        ks = []
        tb = time.perf_counter()
        for _ in range(args.steps):
            ks.append(owner.RenderBatch(rs, wvs, prs, ws, hs))
        el = time.perf_counter() - tb
        return {"mrays_s": round(rays * args.steps / el / 1e6, 1), "kernel_ms": round(float(np.mean(ks)), 3),
                "kernel_mrays_s": round(rays / float(np.mean(ks)) / 1e3, 1), "tuned_chunk": chunk, "tuned_blocks": blocks}

    plain = timed_batch()
    single = L.CreateMultiRenderer(L.DEVICE_GPU)
    single.SetPreset(p)
    items = list(zip(rs, wvs, prs, ws, hs))
    for r, wv, pr, w, h in items:  # warm-up of the per-scene launches
        single.RenderBatch([r], wv[None], pr[None], [w], [h])
    seq_ms = []
    ts = time.perf_counter()
    for _ in range(args.steps):
        seq_ms.append(sum(single.RenderBatch([r], wv[None], pr[None], [w], [h]) for r, wv, pr, w, h in items))
    seq_el = time.perf_counter() - ts
    single.close()
    info = {"scenes": len(rs), "rays": rays, "min_rays": int(min(w * h for w, h in zip(ws, hs))),
            "max_rays": int(max(w * h for w, h in zip(ws, hs))), "bricks_total": int(sum(bricks)),
            "min_bricks": int(min(bricks)), "max_bricks": int(max(bricks)), "scene_build_s": round(build_s, 1),
            "one_launch": plain,
            "launch_per_scene": {"mrays_s": round(rays * args.steps / seq_el / 1e6, 1),
                                 "kernel_ms_sum": round(float(np.mean(seq_ms)), 3)}}
    log(f"batch: {len(rs)} scenes, {rays / 1e6:.1f} Mrays, one launch {plain['kernel_ms']} ms ({plain['mrays_s']} Mrays/s wall, "
        f"chunk {plain['tuned_chunk']}, {plain['tuned_blocks']} blocks); per-scene launches {info['launch_per_scene']['kernel_ms_sum']} "
        f"ms of kernels ({info['launch_per_scene']['mrays_s']} Mrays/s wall)")
    if mlp_params is not None:
        owner.MlpSetParams(mlp_params)
        owner.SetSeeding(True, 0.9)
        info["one_launch_seeded"] = timed_batch()
        owner.SetSeeding(False)
        log(f"batch seeded: {info['one_launch_seeded']['kernel_ms']} ms ({info['one_launch_seeded']['mrays_s']} Mrays/s wall)")
    for r in rs + [owner]:
        r.close()
    return info


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-dr", action="store_true", help="skip the DR leg")
    ap.add_argument("--no-mlp", action="store_true", help="skip the learned-seeder leg")
    ap.add_argument("--lod", type=int, default=128,
                    help="brick LOD of the plain workload's scene (128: 512^3 effective, the headline workload)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched many-scene leg")
    ap.add_argument("--no-shade", action="store_true", help="skip the Phong-with-shadow-rays leg")
    ap.add_argument("--no-build", action="store_true", help="skip the SBS-builder leg")
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--mlp-scenes", type=int, default=32)
    ap.add_argument("--mlp-events", type=int, default=16_000_000)
    ap.add_argument("--mlp-epochs", type=int, default=3)
    ap.add_argument("--calib-copy-mib", type=int, default=0, help="append a device copy of this size (HBM counter calibration)")
    ap.add_argument("--reduce-counters", nargs="+", metavar="DIR",
                    help="reduce rocprofv3 --pmc csv passes found under DIR(s) instead of running (see reduce_counters)")
    ap.add_argument("--kernel", default="k_render<false, false, 2, false>", help="kernel-name substring for --reduce-counters")
    ap.add_argument("--out", default=None, help="output JSON of --reduce-counters")
    args = ap.parse_args()
    if args.reduce_counters:
        reduce_counters(args.reduce_counters, args.kernel, N_VIEWS * W * H, args.calib_copy_mib << 20,
                        args.out or os.path.join(ROOT, "profiles", f"k_render_ico{4 * args.lod}_hbm.json"))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}): the library renders on a single MI355X "
              "and sharding across GPUs is not built.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import litert_amd.api as L

    t0 = time.time()
    sbs = plain_scene(L, args.lod)
    log(f"scene: {sbs.nodes.shape[0]} bricks, {sbs.values.size * 4 / 1e6:.1f} MB values, built in {time.time() - t0:.1f}s")
    r = L.CreateMultiRenderer(L.DEVICE_GPU)
    r.SetScene(sbs)
    pairs, depth = r.GetBVH()
    log(f"bvh: {pairs.shape[0]} pairs, depth {depth}")
    p = L.getDefaultPreset()
    p.sdf_node_intersect = L.SDF_OCTREE_NODE_INTERSECT_NEWTON
    p.render_mode = L.MULTI_RENDER_MODE_NORMAL
    r.SetPreset(p)
    cams = [L.orbit_camera(CAM_SEED, i) for i in range(N_VIEWS)]
    wvs = np.stack([c[0] for c in cams])
    prs = np.stack([c[1] for c in cams])

    for _ in range(args.warmup):
        r.RenderViewsDevice(wvs, prs, W, H, 0, timed=True)
    r.Sync()
    kernel_ms = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        kernel_ms.append(r.RenderViewsDevice(wvs, prs, W, H, 0, timed=True))
    r.Sync()
    elapsed = time.perf_counter() - t_start
    rays = N_VIEWS * W * H
    mrays = rays * args.steps / elapsed / 1e6
    avg_kernel_ms = float(np.mean(kernel_ms))
    log(f"plain: {mrays:.1f} Mrays/s wall, kernel {avg_kernel_ms:.3f} ms/launch ({rays / avg_kernel_ms / 1e3:.1f} Mrays/s)")

    counts = r.CountWork(wvs, prs, W, H)
    bpv = sbs.header.bytes_per_value
    alg_bytes = 64 * int(counts[0]) + 16 * int(counts[1]) + 8 * bpv * int(counts[2]) + 16 * rays
    achieved = alg_bytes / (avg_kernel_ms * 1e-3) / 1e9
    log(f"work/ray: pairs {counts[0] / rays:.3f} bricks {counts[1] / rays:.4f} voxels {counts[2] / rays:.4f}; "
        f"algorithmic {alg_bytes / 1e9:.3f} GB/launch -> {achieved:.0f} GB/s")

    shade_info = None
    if not args.no_shade:
        # Phong shading of the same views: a shadow ray toward the direct light
        # from every hit, RGBA out (eye_ray.cpp:426-478)
        hits = 0
        for v in range(N_VIEWS):
            r.RenderViewsDevice(wvs[v:v + 1], prs[v:v + 1], W, H, 0, timed=True)
            hits += int(np.count_nonzero(r.DownloadView(0, W, H)[2] != 0xFFFFFFFF))
        ps = L.getDefaultPreset()
        ps.sdf_node_intersect = L.SDF_OCTREE_NODE_INTERSECT_NEWTON
        ps.render_mode = L.MULTI_RENDER_MODE_PHONG_NO_TEX
        r.SetPreset(ps)
        r.RenderViewsDevice(wvs, prs, W, H, 1, timed=True)
        sms = [r.RenderViewsDevice(wvs, prs, W, H, 1, timed=True) for _ in range(5)]
        r.SetPreset(p)
        sms = float(np.mean(sms))
        shade_info = {"render_mode": "PHONG_NO_TEX", "kernel_ms": round(sms, 3), "primary_rays": rays, "shadow_rays": hits,
                      "mrays_s_primary_plus_shadow": round((rays + hits) / sms / 1e3, 1)}
        log(f"phong + shadow rays: {sms:.3f} ms/launch, {rays} primary + {hits} shadow rays -> "
            f"{(rays + hits) / sms / 1e3:.1f} Mrays/s")

    build_info = None
    if not args.no_build:
        # the plain scene built from its mesh on the GPU and on the host
        mv, mt = L.mesh_displaced_icosphere(5 if args.lod <= 128 else 6, 0.8, 0.05, 4, SCENE_SEED)
        kw = dict(lod=args.lod, brick_size=4, brick_pad=0, bytes_per_value=1, voxel_diag_scale=True)
        L.sbs_from_mesh(mv, mt, renderer=r, **kw)  # warm-up (module load)
        dms = []
        for _ in range(3):
            dsbs, ms = L.sbs_from_mesh(mv, mt, renderer=r, **kw)
            dms.append(ms)
        hsbs, hms = L.sbs_from_mesh(mv, mt, **kw)
        same = bool(np.array_equal(dsbs.nodes, hsbs.nodes) and np.array_equal(dsbs.values, hsbs.values))
        build_info = {"mesh_triangles": int(mt.shape[0]), "bricks": int(dsbs.nodes.shape[0]), "device_ms": round(min(dms), 1),
                      "host_ms": round(hms, 1), "host_threads": min(len(os.sched_getaffinity(0)), int(os.environ.get("OMP_NUM_THREADS", "0")) or 1 << 30),
                      "identical": same}
        log(f"sbs build from mesh ({mt.shape[0]} triangles, {dsbs.nodes.shape[0]} bricks): device {min(dms):.1f} ms, "
            f"host {hms:.1f} ms, identical {same}")
        r.SetScene(sbs)

    mlp_mrays, mlp_info, mlp_params = None, None, None
    if not args.no_mlp:
        rm = L.CreateMultiRenderer(L.DEVICE_GPU)
        mlp_mrays, mlp_info = mlp_leg(L, np, rm, wvs, prs, args)
        mlp_params = rm.MlpGetParams()
        rm.close()

    batch_info = None if args.no_batch else batch_leg(L, np, args, mlp_params)

    dr_mrays = None
    dr_info = None
    if not args.no_dr:
        tgt, ini, dcams, refs, dr = dr_problem(L, np)
        g = L.MultiRendererDR()
        g.SetScene(ini)
        dwv = np.stack([c[0] for c in dcams])
        dpr = np.stack([c[1] for c in dcams])
        g.SetReference(refs, dwv, dpr)
        g.RenderDR(dr, 0, seed=1)  # warm-up
        dr.opt_iterations = 200  # one view per Adam step (MultiRendererDR.h:40)
        dr.image_batch_size = 1
        work = L.SdfSBS(ini.header, ini.nodes, ini.values, ini.values_f.copy())
        losses, total_ms = g.OptimizeFixedStructure(dr, work, seed=1)
        dres = refs.shape[1]
        dr_mrays = dres * dres * dr.spp * dr.opt_iterations / (total_ms * 1e-3) / 1e6
        view_ms = [g.RenderDR(dr, v, seed=1)[3] for v in range(4)]
        bp, bh = g.LastStats()
        dr_info = {"views": int(refs.shape[0]), "res": int(dres), "border_spp": int(dr.border_spp), "iterations": int(dr.opt_iterations),
                   "ms_per_iteration": round(total_ms / dr.opt_iterations, 3), "renderdr_kernel_ms": round(float(np.mean(view_ms)), 3),
                   "last_view_border_pixels": int(bp),
                   "loss_mean_first_last_16_steps": [float(np.mean(losses[:16])), float(np.mean(losses[-16:]))]}
        log(f"dr: {dr_mrays:.1f} Mrays/s (fwd+bwd+Adam, {total_ms / dr.opt_iterations:.2f} ms/iter), RenderDR kernels "
            f"{np.mean(view_ms):.2f} ms/view, border pixels {bp}")
        # image_batch_size 4 (MultiRendererDR.cpp:395-476): four images per Adam
        # step, the gradient of the last; the other three ride in its primary launch
        dr.opt_iterations = 50
        dr.image_batch_size = 4
        work = L.SdfSBS(ini.header, ini.nodes, ini.values, ini.values_f.copy())
        _, b4_ms = g.OptimizeFixedStructure(dr, work, seed=1)
        dr.image_batch_size = 1
        dr_info["batch4_ms_per_iteration"] = round(b4_ms / 50, 3)
        dr_info["batch4_mrays_s"] = round(dres * dres * 4 * 50 / (b4_ms * 1e-3) / 1e6, 1)
        log(f"dr batch 4: {b4_ms / 50:.3f} ms/iter ({dr_info['batch4_mrays_s']} Mrays/s primary rays of all four images)")
        # SVM border sampling (MultiRendererDR.cpp:1651-1931) on the same views
        # and parameters: per border pixel 64 SVM samples, the perceptron fit
        # and the stripe's border rays, against random sampling above
        dr.dr_border_sampling = L.DR_BORDER_SAMPLING_SVM
        g.RenderDR(dr, 0, seed=1)  # warm-up
        svm_ms = [g.RenderDR(dr, v, seed=1)[3] for v in range(4)]
        sbp, sbh = g.LastStats()
        dr.dr_border_sampling = L.DR_BORDER_SAMPLING_RANDOM
        dr_info["svm_renderdr_kernel_ms"] = round(float(np.mean(svm_ms)), 3)
        dr_info["svm_last_view_border_hits"] = int(sbh)
        dr_info["random_last_view_border_hits"] = int(bh)
        log(f"dr svm: RenderDR kernels {np.mean(svm_ms):.2f} ms/view (random sampling {np.mean(view_ms):.2f}), border hits "
            f"{sbh} vs {bh} on the last view")

    cpu = None
    if not args.no_cpu:
        import oracle
        # the oracle's CPU-baseline build (-O3, AVX2/FMA, no FP contraction);
        # protocol: one warm-up pass, then the median of 5 timed passes
        sc = oracle.Scene(sbs, pairs, baseline=True)
        nthreads = min(len(os.sched_getaffinity(0)), int(os.environ.get("OMP_NUM_THREADS", "0")) or 1 << 30)
        model = cpu_model()

        def full_pass(threads):
            t = time.perf_counter()
            for v in range(N_VIEWS):
                sc.render(wvs[v], prs[v], p, W, H, precision=32, threads=threads)
            return time.perf_counter() - t

        full_pass(nthreads)  # warm-up
        times = sorted(full_pass(nthreads) for _ in range(5))
        cpu_mrays = rays / times[2] / 1e6
        one = []
        for _ in range(3):
            t1 = time.perf_counter()
            sc.render(wvs[0], prs[0], p, 1024, 1024, precision=32, threads=1)
            one.append(time.perf_counter() - t1)
        one_thread = 1024 * 1024 / sorted(one)[1] / 1e6
        cpu = {"value": round(cpu_mrays, 2), "unit": "Mrays/s", "cores": nthreads, "kind": "port",
               "sample": f"the full workload (8 views x {W}x{H}, same scene, cameras, Newton preset): median of 5 passes after one "
                         f"warm-up ({', '.join(f'{t:.2f}' for t in times)} s), FP32 oracle built -O3 -march=x86-64-v3, OpenMP "
                         f"{nthreads} threads on {model}; 1 thread: {one_thread:.2f} Mrays/s (median of 3 on one 1024x1024 view)"}
        log(f"cpu baseline: {cpu_mrays:.2f} Mrays/s on {nthreads} threads ({model}), {one_thread:.2f} on 1 thread")
        if dr_info is not None:
            tgt, ini, dcams, refs, dr = dr_problem(L, np, views=1)
            ip, _ = L.build_bvh(ini)
            si = oracle.Scene(ini, ip, baseline=True)
            ref = refs[0].copy()
            ref[..., 3] = np.linalg.norm(ref[..., :3], axis=-1) >= 0.001

            def dr_pass():
                t2 = time.perf_counter()
                si.dr_render_grad(dr, dcams[0][0], dcams[0][1], refs.shape[2], refs.shape[1], ref, ini.values_f.size, seed=1,
                                  precision=32, threads=nthreads)
                return time.perf_counter() - t2

            dr_pass()  # warm-up
            dts = sorted(dr_pass() for _ in range(5))
            dt = dts[2]
            dr_info["cpu_baseline_mrays_s"] = round(refs.shape[1] * refs.shape[2] / dt / 1e6, 3)
            dr_info["cpu_baseline_sample"] = (f"one RenderDR view (fwd+bwd incl. border rays), FP32 oracle baseline build, {nthreads} "
                                              f"threads, median of 5 after a warm-up: {dt:.3f}s")
            log(f"dr cpu baseline: {dr_info['cpu_baseline_mrays_s']} Mrays/s")

    if args.calib_copy_mib:
        # plain device-to-device copy through the HIP runtime's blit kernel:
        # n bytes read and n bytes written, the scale for HBM byte counters
        import ctypes as C
        hip = C.CDLL("libamdhip64.so.7")
        n = C.c_size_t(args.calib_copy_mib << 20)
        a, b = C.c_void_p(), C.c_void_p()
        for call in (lambda: hip.hipMalloc(C.byref(a), n), lambda: hip.hipMalloc(C.byref(b), n),
                     lambda: hip.hipMemset(a, 1, n), lambda: hip.hipMemcpy(b, a, n, 3), lambda: hip.hipDeviceSynchronize(),
                     lambda: hip.hipFree(a), lambda: hip.hipFree(b)):
            rc = call()
            if rc != 0:
                raise RuntimeError(f"HIP call failed with {rc} in the calibration copy")
        log(f"calibration copy: {n.value} bytes read + written")

    counters = measured_counters(args.lod)
    out = {
        "metric": "primary Mrays/s on SBS (plain vs MLP-seeded); DR fwd+bwd Mrays/s",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded displaced icosphere SBS, seeded orbit cameras)",
        "config": {"workload": f"displaced icosphere SBS {4 * args.lod}^3 effective (LOD {args.lod} x 4^3 bricks, 8-bit), "
                               "8 views x 2048x2048 primary rays, Newton",
                   "bricks": int(sbs.nodes.shape[0]), "rays_per_step": rays, "parallelism": "single GPU"},
        "kernel_ms_per_step": round(avg_kernel_ms, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": counters.get("traffic_bytes"),
                     "algorithmic_bytes_per_launch": alg_bytes,
                     # what limits the kernel instead: the VALU pipes issue on every
                     # cycle (valu_issue_frac ~1) with part of each wave's lanes active
                     "limiter": "valu_issue" if counters.get("valu_issue_frac", 0) >= 0.9 else None,
                     "valu_issue_frac": counters.get("valu_issue_frac"),
                     "valu_lane_util": counters.get("valu_lane_utilisation")},
        "cpu_baseline": cpu,
        "mlp_seeded_mrays_s": mlp_mrays,
        "mlp": mlp_info,
        "batch": batch_info,
        "phong_shadow": shade_info,
        "sbs_build": build_info,
        "dr_fwd_bwd_mrays_s": None if dr_mrays is None else round(dr_mrays, 2),
        "dr": dr_info,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
