"""Frame-loop benchmark: frames/s of integrate + raycast + 3-level ICP on one MI355X.

Workload: a 512^3 TSDF at 2 mm over a 1.024 m cube, 640x480 synthetic uint16 depth
frames (sphere r = 0.3 m before the plane z = 1.2 m, 10 seeded boxes, orbit 2 deg /
frame with 5 mm / 0.5 deg jitter, Gaussian depth noise 0.0012 z^2), ICP {10, 5, 4}
iterations. Each step is one frame of the loop in the reference pipeline's order
(live pyramid -> ICP against the model raycast at the previous pose -> integrate ->
raycast). Frames are uploaded to HBM before the timed region; the tracked pose stays
on the device and is copied out asynchronously, so the timed region is GPU work only.

`value` is the frames/s of the FP16 + UInt16 TSDF (4 B/voxel); the plain FP32 TSDF
(8 B/voxel, the bit-exact parity kernels) runs the same frames and is reported beside
it. `roofline` is for the dominant kernel of the FP16 form (the one with the most
time per frame), from HIP events on its launch stream; `cpu_baseline` times the CPU
oracle (FP32 restatement of the reference CPU path with the reference's 20 B AoS
voxel) on the first frames of the same sequence, and counts the voxels the raycast
reads there (the raycast's algorithmic bytes).

Beside the frame loop (skipped with --no-extras): the 1024^3 / 1280x960 form (20
frames each of FP16 and FP32), the raycast-depth loss and its adjoint on the 512^3
FP32 model (640x480 rays, step 0.5 voxel), and the raw-Kinect depth conditioning chain
on one 640x480 frame. `accuracy_vs_fp64` is the FP16 form's error statement against
the FP64 oracle as recorded in profiles/accuracy_512.json (the benchmark does not run
the oracle outside its CPU-baseline leg).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "frames/s (integrate+raycast+ICP) at 512³, FP16 vs FP32 TSDF; % HBM peak"


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=97)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2_512")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-frames", type=int, default=8)
    ap.add_argument("--profile-frames", type=int, default=24)
    ap.add_argument("--only", choices=["both", "f16", "f32"], default="both")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the legs beside the frame loop (1024^3 form, adjoint, depth conditioning)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy kernel of MIB MiB 5 times (the known-size "
                         "reference for HBM byte counters) and exit")
    return ap.parse_args()


def copy_calibration(mib):
    import torch

    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    assert bool((dst[-4:] == src[-4:]).all())
    log("copied %d MiB x 5" % mib)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_form(frames_dev, nframes, pose0, params, fmt, W, H, K, warmup, steps, keep=False):
    """Times `steps` frames after `warmup`; with keep, also returns the open loop (its
    volume holds the fused model, its tracker the last live frame)."""
    from azu_amd import native as N
    from azu_amd.pipeline import FrameLoop

    loop = FrameLoop(params, fmt, W, H, K)
    log("verified fixed-divisor division (voxel size, truncation):", loop.vol.division_check())
    loop.reset(pose0)
    for f in range(warmup):
        loop.frame(frames_dev[f])
    N.synchronize()
    t0 = time.perf_counter()
    for f in range(warmup, warmup + steps):
        loop.frame(frames_dev[f])
    N.synchronize()
    dt = time.perf_counter() - t0
    poses = loop.poses()
    if keep:
        return dt, poses, loop
    loop.close()
    return dt, poses


def adjoint_leg(loop, pose, K, W, H, reps=5):
    """The raycast-depth loss and its adjoint on the FP32 model the loop fused (512^3),
    640 x 480 rays, step 0.5 voxel, observed depth = the last live frame: event-timed
    forward (loss only) and forward + backward (pose gradient and the TSDF gradient
    scattered into a res^3 float buffer)."""
    from azu_amd import native as N
    from azu_amd.fusion import pose_cm

    d_obs = loop.tr.level(0)[0]
    res = loop.vol.res
    grad = N.DeviceBuffer(res ** 3 * 4)  # contents irrelevant to the timing (+= into it)
    Kc = N.f32(K)
    T = N.f32(pose_cm(pose))
    out = (ctypes.c_double * 8)()

    def call(backward):
        N.check(N.lib.azu_raycast_depth_loss(loop.vol.h, N.fptr(T), N.fptr(Kc), W, H, d_obs, 0.5,
                                             1 if backward else 0, grad.ptr if backward else None, None, out,
                                             None))

    call(False)
    call(True)
    N.profile_reset()
    N.profile_enable(True)
    for _ in range(reps):
        call(False)
    for _ in range(reps):
        call(True)
    prof = N.profile_read()
    N.profile_enable(False)
    grad.free()
    fwd, bwd = prof.get("depth_loss_fwd"), prof.get("depth_loss_bwd")
    return {"volume": "%d^3 fp32" % res, "rays": "%dx%d" % (W, H), "step_voxels": 0.5,
            "forward_us": round(1000 * fwd[1] / fwd[0], 1) if fwd else None,
            "forward_backward_us": round(1000 * bwd[1] / bwd[0], 1) if bwd else None,
            "pixels": int(out[7]), "loss": float(out[0])}


def conditioning_leg(frame_mm, reps=50):
    """The raw-Kinect depth conditioning chain (median, hole fill, guided filter with a
    guidance luma image, temporal EMA) on one 640 x 480 frame of raw codes made from the
    sequence's first frame, per-stage event times averaged over reps frames."""
    from azu_amd import native as N
    from azu_amd.conditioner import SignalConditioner

    H, W = frame_mm.shape
    m = frame_mm.astype(np.float64) * 1e-3
    with np.errstate(divide="ignore"):
        code = (1.0 / np.where(m > 0, m, np.inf) - 3.3309495161) / -0.0030711016
    raw = np.where(frame_mm > 0, np.clip(np.rint(code), 1, 2046), 0).astype(np.uint16)
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float32)
    luma = np.clip(0.4 + 0.3 * xx / W + 0.1 * np.sin(yy / 23.0), 0.0, 1.0).astype(np.float32)
    sc_ = SignalConditioner(W, H)
    d_in, d_out, d_luma = N.device_array(raw), N.DeviceBuffer(raw.nbytes), N.device_array(luma)
    for _ in range(3):
        sc_.process_device(d_in.ptr, d_out.ptr, d_luma.ptr)
    N.synchronize()
    N.profile_reset()
    N.profile_enable(True)
    for _ in range(reps):
        sc_.process_device(d_in.ptr, d_out.ptr, d_luma.ptr)
    prof = N.profile_read()
    N.profile_enable(False)
    N.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        sc_.process_device(d_in.ptr, d_out.ptr, d_luma.ptr)
    N.synchronize()
    wall = (time.perf_counter() - t0) / reps
    sc_.close()
    stages = {k: round(1000 * v[1] / v[0], 2) for k, v in prof.items() if k.startswith("cond_")}
    return {"frame": "%dx%d raw codes + luma" % (W, H), "us_per_frame": round(1e6 * wall, 1),
            "stage_us": stages}


def c5_leg(steps=20, warmup=2):
    """The large-volume form: 1024^3 at 1 mm, 1280 x 960 frames (the 640 x 480 sequence
    upsampled 2x bicubic), FP16 + UInt16 and FP32, `steps` timed frames each."""
    from azu_amd import native as N
    from azu_amd import scene as sc

    w = sc.WORKLOADS["c5_1024"]
    params = sc.volume_params("c5_1024")
    frames_mm, poses_true, K = sc.sequence("c5_1024", frames=warmup + steps)
    dev = [N.device_array(f) for f in frames_mm]
    out = {"workload": "1024^3 TSDF at 1 mm, 1280x960 uint16 depth (2x bicubic), %d frames" % (warmup + steps)}
    for name, fmt in (("f16", N.FMT_F16_U16), ("f32", N.FMT_F32)):
        dt, poses = run_form([d.ptr for d in dev], warmup + steps, poses_true[0], params, fmt, w["W"], w["H"], K,
                             warmup, steps)
        terr = np.abs(poses[:, :3, 3] - poses_true[:warmup + steps, :3, 3]).max()
        out[name] = {"fps": round(steps / dt, 2), "ms_per_step": round(1000 * dt / steps, 3),
                     "max_track_err_m": float(terr)}
        log("c5", name, out[name])
    for d in dev:
        d.free()
    out["fp16_over_fp32"] = round(out["f16"]["fps"] / out["f32"]["fps"], 3)
    return out


def accuracy_statement():
    """The FP16 + UInt16 form's error against the FP64 oracle, as measured by
    tests/test_gpu_full_size.py::test_fp16_error_statement and recorded in
    profiles/accuracy_512.json (the oracle is not run by the benchmark)."""
    try:
        with open(os.path.join(ROOT, "profiles", "accuracy_512.json")) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    a, s = rec["fp16_vs_fp64_oracle_10_frames"], rec["fp16_vs_fp32_oracle_10_frames"]
    return {"source": "profiles/accuracy_512.json (%s)" % rec["command"],
            "fp16_vs_fp64": {k: a[k] for k in ("tsdf_rms", "depth_rms_m", "depth_p999_m", "hit_mismatch")},
            "fp16_storage_only_vs_fp32": {k: s[k] for k in ("tsdf_rms", "tsdf_max", "depth_rms_m")}}


def run_form_pcie(frames_mm, pose0, params, fmt, W, H, K, warmup, steps):
    """The same loop with each frame handed over in pinned host memory: one host-to-device
    copy of the uint16 frame per step on the loop's stream, inside the timed region."""
    from azu_amd import native as N
    from azu_amd.pipeline import FrameLoop

    n = warmup + steps
    fb = W * H * 2
    host = ctypes.c_void_p()
    N.check(N.lib.azu_host_alloc(ctypes.byref(host), fb * n))
    frames_c = np.ascontiguousarray(frames_mm[:n])  # referenced until the copy has run
    ctypes.memmove(host.value, frames_c.ctypes.data, fb * n)
    slot = N.DeviceBuffer(fb)
    loop = FrameLoop(params, fmt, W, H, K)
    loop.reset(pose0)
    try:
        for f in range(n):
            if f == warmup:
                N.synchronize()
                t0 = time.perf_counter()
            N.check(N.lib.azu_memcpy_h2d(slot.ptr, host.value + f * fb, fb, None))
            loop.frame(slot.ptr)
        N.synchronize()
        dt = time.perf_counter() - t0
    finally:
        loop.close()
        N.lib.azu_host_free(host.value)
    return steps / dt


def profile_form(frames_dev, pose0, params, fmt, W, H, K, nframes):
    """Per-kernel averages over nframes of a fresh run, HIP events on the launch stream."""
    from azu_amd import native as N
    from azu_amd.pipeline import FrameLoop

    loop = FrameLoop(params, fmt, W, H, K)
    loop.reset(pose0)
    loop.frame(frames_dev[0])
    N.synchronize()
    loop.vol.counters(reset=True)
    work = (ctypes.c_double * 5)()
    N.check(N.lib.azu_tracker_phase_times(loop.tr.h, work, 1))
    N.profile_reset()
    N.profile_enable(True)
    for f in range(1, nframes):
        loop.frame(frames_dev[f])
    N.synchronize()
    prof = N.profile_read()
    N.profile_enable(False)
    upd, _ = loop.vol.counters(reset=True)
    N.check(N.lib.azu_tracker_phase_times(loop.tr.h, work, 1))
    loop.close()
    nf = max(nframes - 1, 1)
    icp = {"iterations_run": work[2] / nf, "pixels": work[3] / nf, "projections": work[4] / nf,
           "sum_us": 1000 * work[0] / nf, "solve_us": 1000 * work[1] / nf}
    return prof, upd / nf, icp


def rooflines(prof, nprof, n_upd, icp, s_v, W, H, res, u_ray=None):
    """Achieved HBM GB/s of each hot kernel from its algorithmic bytes per frame and its
    event-timed duration per frame (equal to bytes per launch / average launch time)."""
    out = {}
    for k, (launches, total_ms) in prof.items():
        ms_frame = total_ms / nprof
        if k.startswith("integrate"):
            # read + write of every updated voxel, plus the float depth frame read once
            b = 2 * s_v * n_upd + 4 * W * H
            # the same time against the bytes of a read + write of the whole volume
            full = 2 * s_v * res ** 3 / (total_ms / launches * 1e-3) / 1e9
            extra = {"updated_voxels": int(n_upd), "full_sweep_equiv_frac": round(full / HBM_PEAK_GBS, 4)}
        elif k.startswith("raycast") and u_ray:
            # the distinct voxels the marches and normal stencils read, once each, plus the
            # vertex and normal image written
            b = s_v * u_ray + 24 * W * H
            extra = {"touched_voxels": int(u_ray)}
        elif k == "icp_iter":
            # live vertex + normal per interior pixel, model vertex + normal per in-bounds
            # projection, and the per-block partials (34 doubles x at most 256 blocks, written
            # and read back; 256 is the level-0 count and an upper bound for the coarser
            # levels' 150 and 38) of every iteration that ran
            b = 24 * icp["pixels"] + 24 * icp["projections"] + 2 * 34 * 8 * 256 * icp["iterations_run"]
            extra = {"iterations_run": round(icp["iterations_run"], 2), "pixels": int(icp["pixels"]),
                     "projections": int(icp["projections"])}
        else:
            continue
        gbs = b / (ms_frame * 1e-3) / 1e9
        out[k] = dict({"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None, "kernel": k,
                       "avg_launch_us": round(1000 * total_ms / launches, 2),
                       "algorithmic_bytes_per_launch": int(b * nprof / launches)}, **extra)
    return out


def measured_traffic(kernel, workload, form):
    """HBM bytes per launch of `kernel` from the rocprofv3 FETCH_SIZE / WRITE_SIZE
    passes recorded in profiles/traffic.json (each counter scaled by what it reads for a
    copy kernel of known size), or None when that kernel and workload were not measured."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    e = rec.get("%s/%s/%s" % (workload, form, kernel))
    return None if e is None else int(e["hbm_bytes_per_launch"])


def cpu_baseline(frames_mm, pose0, params, K, nframes):
    import oracle

    v = oracle.Volume(params["resolution"], params["voxel_size"], params["truncation"], params["max_weight"],
                      params["origin"], params["min_depth"], params["max_depth"], dtype=np.float32)
    t0 = time.perf_counter()
    poses, ms, _ = v.run_sequence(frames_mm[:nframes], K, pose0)
    dt = time.perf_counter() - t0
    H, W = frames_mm.shape[1:]
    # Distinct voxels the last frame's raycast reads (the raycast's algorithmic bytes);
    # counted on the oracle's model, which the FP32 kernels reproduce bit for bit.
    u_ray = int(v.raycast_touched(poses[-1], K, W, H))
    v.close()
    # frame 0 has no ICP; the rate is over frames 1.. (median frame time)
    med = float(np.median(ms[1:, 4])) if nframes > 1 else float(ms[0, 4])
    return {"value": 1000.0 / med, "unit": "frames/s", "cores": int(oracle.lib.orc_num_threads()), "kind": "port",
            "sample": "frames 1..%d of the same %d^3 sequence (%.1f s wall), median frame time %.1f ms "
                      "(pyramid %.1f, icp %.1f, integrate %.1f, raycast %.1f ms)" % (
                          nframes - 1, params["resolution"], dt, med, np.median(ms[1:, 0]), np.median(ms[1:, 1]),
                          np.median(ms[1:, 2]), np.median(ms[1:, 3])), "_u_ray": u_ray}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the frame loop on a single MI355X "
              "(the multi-GPU slab decomposition is a design note, not built); use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    from azu_amd import native as N
    from azu_amd import scene as sc

    w = sc.WORKLOADS[args.workload]
    params = sc.volume_params(args.workload)
    nframes = args.warmup + args.steps
    t0 = time.perf_counter()
    frames_mm, poses_true, K = sc.sequence(args.workload, frames=max(nframes, args.profile_frames))
    log("rendered %d frames in %.1f s" % (len(frames_mm), time.perf_counter() - t0))
    W, H = w["W"], w["H"]
    dev = [N.device_array(f) for f in frames_mm]
    frames_dev = [d.ptr for d in dev]
    pose0 = poses_true[0]
    dev_id, dev_name, hbm = N.device_info()
    log("device", dev_name)

    results = {}
    adjoint = None
    forms = {"f16": N.FMT_F16_U16, "f32": N.FMT_F32}
    for name in (["f16", "f32"] if args.only == "both" else [args.only]):
        keep = name == "f32" and not args.no_extras
        r = run_form(frames_dev, nframes, pose0, params, forms[name], W, H, K, args.warmup, args.steps, keep=keep)
        dt, poses = r[0], r[1]
        if keep:
            adjoint = adjoint_leg(r[2], poses[-1], K, W, H)
            r[2].close()
            log("adjoint", adjoint)
        terr = np.abs(poses[:, :3, 3] - poses_true[:nframes, :3, 3]).max(axis=1)
        results[name] = {"seconds": dt, "fps": args.steps / dt, "ms_per_step": 1000.0 * dt / args.steps,
                         "max_track_err_m": float(terr.max())}
        log(name, "%.1f frames/s, %.3f ms/frame, max |t - t_true| %.2f mm" % (
            results[name]["fps"], results[name]["ms_per_step"], 1000 * terr.max()))

    head = "f16" if "f16" in results else "f32"
    pcie = run_form_pcie(frames_mm, pose0, params, forms[head], W, H, K, args.warmup, args.steps)
    log("%s with per-frame upload from pinned host memory: %.1f frames/s" % (head, pcie))

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(frames_mm, pose0, params, K, args.cpu_frames)
        log("cpu baseline", cpu)
    u_ray = cpu.pop("_u_ray") if cpu else None

    # Per-kernel timing (fresh run, events on the launch stream) for the roofline.
    nprof_frames = max(min(args.profile_frames, len(frames_dev)), 2)
    prof, n_upd, icp = profile_form(frames_dev, pose0, params, forms[head], W, H, K, nprof_frames)
    nprof = nprof_frames - 1
    per_frame = {k: v[1] / nprof for k, v in prof.items()}
    # The dominant kernel among the one-launch-per-frame kernels: the ICP's 20 launches a
    # frame each carry their own event pair, whose recording gaps the per-frame sum of
    # icp_iter includes (its rocprof kernel time per frame is below integrate's).
    single = {k: v for k, v in per_frame.items() if k.startswith(("integrate", "raycast"))}
    dom = max(single or per_frame, key=(single or per_frame).get)
    s_v = 4 if head == "f16" else 8
    roofs = rooflines(prof, nprof, n_upd, icp, s_v, W, H, w["res"], u_ray)
    roof = roofs.get(dom) or {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                              "traffic": None, "kernel": dom,
                              "avg_launch_us": round(1000 * prof[dom][1] / prof[dom][0], 2)}
    traffic = measured_traffic(roof["kernel"], args.workload, head)
    if traffic is not None:
        roof["traffic"] = traffic
    log("icp work per frame:", json.dumps(icp))
    kern = {k: {"launches": int(v[0]), "avg_us": round(1000 * v[1] / v[0], 2),
                "us_per_frame": round(1000 * v[1] / nprof, 2)} for k, v in prof.items()}
    log("kernels:", json.dumps(kern))

    extras = {}
    if not args.no_extras:
        extras["depth_conditioning"] = conditioning_leg(frames_mm[0])
        log("conditioning", extras["depth_conditioning"])
        for d in dev:
            d.free()
        extras["c5_1024"] = c5_leg()

    r = results[head]
    line = {
        "metric": METRIC,
        "value": round(r["fps"], 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "%d^3 TSDF at %g mm, %dx%d uint16 depth%s, sphere+plane%s, orbit%s, "
                        "integrate+raycast+ICP{10,5,4}" % (
                            w["res"], 1000 * w["vs"], W, H, " (2x bicubic upsampled)" if W != 640 else "",
                            "+%d boxes" % w["boxes"] if w["boxes"] else "", "+jitter" if w["jitter"] else ""),
            "tsdf_storage": "fp16+u16 (4 B/voxel)" if head == "f16" else "fp32+fp32 (8 B/voxel)",
            "frames": nframes,
            "full_sweep_bytes_per_integrate": 2 * (4 if head == "f16" else 8) * w["res"] ** 3,
            "device": dev_name,
        },
        "fp32_tsdf": ({"value": round(results["f32"]["fps"], 2), "ms_per_step": round(results["f32"]["ms_per_step"], 4),
                       "max_track_err_m": results["f32"]["max_track_err_m"]} if "f32" in results else None),
        "fp16_over_fp32": (round(results["f16"]["fps"] / results["f32"]["fps"], 3)
                           if "f16" in results and "f32" in results else None),
        "max_track_err_m": r["max_track_err_m"],
        "pcie_inclusive_fps": round(pcie, 2),
        "roofline": roof,
        "roofline_by_kernel": {k: {kk: v[kk] for kk in ("achieved", "frac", "avg_launch_us", "full_sweep_equiv_frac",
                                                        "algorithmic_bytes_per_launch")
                                   if kk in v} for k, v in roofs.items()},
        "accuracy_vs_fp64": accuracy_statement(),
        "c5_1024": extras.get("c5_1024"),
        "adjoint_c4": adjoint,
        "depth_conditioning": extras.get("depth_conditioning"),
        "kernels_us_per_frame": {k: v["us_per_frame"] for k, v in kern.items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
