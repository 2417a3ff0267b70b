#!/usr/bin/env python3
"""Headline benchmark: ptnee camera samples per second at 720p on one MI355X.

Workload (one step): a 1280x720 frame at 64 samples per pixel, max depth 10, Russian roulette
(prr 0.8) from vertex 3, on the procedural scene of a 512x512-cell heightfield, 2000 80-triangle
icospheres with GGX-reflect materials, four emissive quads and a 1024x512 sky (684,296
triangles).  The scene is built and resident in HBM before the timed region; a step is one
nr_render_ptnee call (all wavefront batches of the frame, FP64 accumulation included).

The inverse-rendering iteration (render + loss + path-replay adjoint + Adam on the 256x256
64-material box room) is timed after the frames and reported as "adjoint"; the same frame with
sky importance sampling as "env_sampling" (time, variance, equal-error gain); the time to RMSE
1e-2 of plain ptnee and of the radiance cache at 1920x1080, 1 spp per frame, as "nrc"; the
reference's own box room at 320x240 x 16 spp (GPU wavefront, GPU path kernel, CPU oracle) as
"reference_scene".

Printed JSON line: value = Msamples/s over the timed steps, the traversal roofline (HIP events
around every closest-hit launch; algorithmic bytes from the kernel's own node / triangle
counters) and the CPU baseline (the test oracle's restatement of the reference estimator timed
on a bounded slice of the same frame with all host threads).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# VALU issue peak: a wave64 instruction occupies a 16-lane SIMD for 4 cycles, so 256 CUs x 4 SIMDs x
# 16 lanes x 2.4 GHz = 39.3 T lane-operations per second
VALU_PEAK_TLANEOPS = 256 * 4 * 16 * 2.4e9 / 1e12
# algorithmic bytes of one closest-hit ray: 128 B per 4-wide node fetch (four child boxes and
# references), 48 B per triangle test, 32 B ray in, 16 B hit out
B_NODE, B_TRI, B_RAY = 128, 48, 48
# algorithmic FP32 operations of the reference's tests: a slab test (bvhnode.cpp:35-48) is 6 sub +
# 6 div + 12 min / max / compare = 24, so a 4-wide node fetch (four child boxes) is 96; a triangle
# test (triangle.cpp:27-33) is 9 sub + 12 for the two crosses + 4 dots + 3 div = 44 (each one
# lane-operation, whatever the instructions the kernel spends on it)
OPS_NODE, OPS_TRI = 4 * 24, 44


# crops of the headline frame checked against the FP64 oracle (x0, y0, x1, y1)
MRE_CROPS = [(100, 60, 148, 108), (600, 300, 648, 348), (1100, 500, 1148, 548), (640, 80, 688, 128)]


def cpu_baseline(sc, spp_frame, gpu_mean=None, target_s=12.0):
    """Times the oracle (CPU restatement of RendererPTLS, OpenMP over 8x8 tiles, all host
    threads) on the same frame: one calibration pass at 1 spp, then k spp sized to ~target_s.
    With the GPU frame's per-pixel mean (same seed), also its mean relative error against the FP64
    oracle on four 48 x 48 crops at the frame's spp: mean |gpu - fp64| / (|fp64| + 1e-2), the same
    random numbers (the oracle consumes the full frame's Philox streams on a crop)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    from nuray_amd.scenes import camera_array

    threads = len(os.sched_getaffinity(0))
    W, H = sc["width"], sc["height"]
    o = po.Scene(sc)
    cam = camera_array(sc["camera"])
    kw = dict(maxdepth=10, rrdepth=3, nthreads=threads, seed=7)
    t0 = time.perf_counter()
    o.render(cam, W, H, 1, **kw)
    dt = time.perf_counter() - t0
    k = int(np.clip(round(target_s / max(dt, 1e-3)), 1, spp_frame))
    t0 = time.perf_counter()
    o.render(cam, W, H, k, spp0=1, **kw)
    dt = time.perf_counter() - t0
    samples = W * H * k
    # one host thread on every 8th row of the frame (a bounded sample spread over the whole image)
    one = 0
    t1 = time.perf_counter()
    for y in range(0, H, 8):
        o.render(cam, W, H, 2, spp0=1 + k, rect=(0, y, W, y + 1), **dict(kw, nthreads=1))
        one += W * 2
    dt1 = time.perf_counter() - t1
    mre = None
    if gpu_mean is not None:
        sums = (np.zeros(W * H * 3), np.zeros(W * H * 3))
        for rect in MRE_CROPS:
            o.render(cam, W, H, spp_frame, rect=rect, sums=sums, fp64=True, **kw)
        ref = sums[0].reshape(H, W, 3) / spp_frame
        g = np.concatenate([gpu_mean[y0:y1, x0:x1].reshape(-1) for x0, y0, x1, y1 in MRE_CROPS])
        r = np.concatenate([ref[y0:y1, x0:x1].reshape(-1) for x0, y0, x1, y1 in MRE_CROPS])
        mre = {"value": float((np.abs(g - r) / (np.abs(r) + 1e-2)).mean()), "crops": MRE_CROPS,
               "spp": spp_frame, "definition": "mean over crop pixel channels of |gpu - fp64| / (|fp64| + 1e-2), "
                                              "same seed (same Philox numbers)"}
    return {"value": samples / dt / 1e6, "unit": "Msamples/s", "cores": threads, "kind": "port",
            "mre_fp64": mre,
            "sample": f"the same 1280x720 frame at {k} spp (samples 1..{k}; {samples} camera samples in {dt:.1f} s)",
            "cpu_model": _cpu_model(), "one_thread_value": one / dt1 / 1e6,
            "one_thread_sample": f"every 8th row of the frame at 2 spp ({one} camera samples in {dt1:.1f} s)"}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def adjoint_iteration(nr, scenes, iters=20, warmup=5):
    """The inverse-rendering iteration on the 64-material box room: upload theta, 256x256 render
    at 32 spp (one lane per path: the same image as the wavefront, bit for bit, without a launch
    per bounce), L2 loss against a 1024-spp target, path-replay adjoint to the 256 parameters,
    Adam step on the host.  Returns ms per iteration and its parts (the wavefront render of the
    same frame timed beside it)."""
    sc, theta_star, theta0 = scenes.adjoint_problem(width=256, height=256)
    W, H, M, cam = 256, 256, len(theta0), sc["camera"]
    dev = nr.Scene(sc)
    dev.set_params(theta_star)
    tgt = dev.render(cam, W, H, 1024, seed=1234)["rgb"]
    emit = sc["mats"][:, 0] == scenes.LIGHT
    theta = theta0.astype(np.float64)
    m = np.zeros_like(theta)
    v = np.zeros_like(theta)

    def step(it):
        nonlocal theta, m, v
        dev.set_params(theta.astype(np.float32))
        dev.render(cam, W, H, 32, seed=it, fetch=False, path_kernel=True)
        dev.loss_l2(tgt)
        g = dev.adjoint(cam, W, H, 32, M, seed=it)
        g[emit] = 0
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        upd = 1e-2 * (m / (1 - 0.9 ** it)) / (np.sqrt(v / (1 - 0.999 ** it)) + 1e-8)
        theta = np.where(emit[:, None], theta, np.clip(theta - upd, [1e-3] * 3 + [0.02], 1.0))
    for it in range(1, warmup + 1):
        step(it)
    nr.synchronize()
    t0 = time.perf_counter()
    for it in range(warmup + 1, warmup + iters + 1):
        step(it)
    nr.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / iters
    # parts, timed separately after the loop
    nr.synchronize()
    t0 = time.perf_counter()
    for it in range(iters):
        dev.render(cam, W, H, 32, seed=it, fetch=False, path_kernel=True)
    nr.synchronize()
    render_ms = (time.perf_counter() - t0) * 1e3 / iters
    # the replay of the last render's samples (its per-sample radiance and vertex records), as in
    # the iteration
    dev.loss_l2(tgt)
    t0 = time.perf_counter()
    for it in range(iters):
        dev.adjoint(cam, W, H, 32, M, seed=iters - 1)
    nr.synchronize()
    adjoint_ms = (time.perf_counter() - t0) * 1e3 / iters
    t0 = time.perf_counter()
    for it in range(iters):
        dev.render(cam, W, H, 32, seed=it, fetch=False)
    nr.synchronize()
    wavefront_ms = (time.perf_counter() - t0) * 1e3 / iters
    return {"ms_per_iter": ms, "render_ms": render_ms, "adjoint_ms": adjoint_ms, "wavefront_render_ms": wavefront_ms,
            "workload": "box room, 64 GGX-reflect materials (256 parameters), 256x256 at 32 spp per "
                        "iteration, L2 loss, path-replay adjoint, Adam lr 1e-2"}


def nrc_time_to_rmse(nr, scenes, eps=1e-2, ref_spp=4096, warm=256, k_small=16, k_large=512):
    """Time to RMSE `eps` against a high-spp reference, plain ptnee vs the radiance cache, on the
    1920x1080 heightfield + spheres scene at 1 spp per frame (maxdepth 10, RR from vertex 3).

    Reference: ptnee at ref_spp spp (FP64 sums); the reference's own variance (from its sums of
    squares) is subtracted from every measured MSE, so the figures estimate the error against the
    expectation.  Plain PT is unbiased: MSE(K frames) = s2 / K, s2 measured on a 256-frame
    accumulation, time = frame_ms * s2 / eps^2.  Radiance cache: `warm` frames of online training
    from a fresh network (timed, not accumulated), then frames that keep training accumulated from
    scratch; MSE(K) = b2 + s2 / K fitted at K = k_small and k_large, time = warm-up +
    frame_ms * s2 / (eps^2 - b2) (never, if b2 >= eps^2)."""
    W, H = 1920, 1080
    sc = scenes.terrain_spheres(width=W, height=H)
    cam = sc["camera"]
    kw = dict(maxdepth=10, rrdepth=3)
    dev = nr.Scene(sc)
    t0 = time.perf_counter()
    r = dev.render(cam, W, H, ref_spp, seed=999, **kw)
    ref = r["sum"] / ref_spp
    var_ref = float(np.maximum(r["sum2"] / ref_spp - ref ** 2, 0.0).mean() / ref_spp)
    ref_s = time.perf_counter() - t0

    def mse():
        return float(((dev.fetch(W, H)["sum"] / dev_spp[0] - ref) ** 2).mean()) - var_ref
    dev_spp = [1]

    def legs(env_is):
        # plain ptnee
        K = 256
        nr.synchronize()
        t0 = time.perf_counter()
        for k in range(K):
            dev.render(cam, W, H, 1, seed=5, spp_offset=k, accumulate=k > 0, fetch=False, env_is=env_is, **kw)
        nr.synchronize()
        pt_ms = (time.perf_counter() - t0) * 1e3 / K
        dev_spp[0] = K
        pt_s2 = mse() * K
        pt_time = pt_ms * 1e-3 * pt_s2 / eps ** 2
        # radiance cache
        cache = nr.Nrc(seed=1)
        nr.synchronize()
        t0 = time.perf_counter()
        for k in range(warm):
            dev.render_nrc(cache, cam, W, H, spp=1, seed=3, spp_offset=k, env_is=env_is, **kw)
        nr.synchronize()
        warm_s = time.perf_counter() - t0
        fit = {}
        frame = 0
        for kk in (k_small, k_large):
            nr.synchronize()
            t0 = time.perf_counter()
            for k in range(kk):
                st = dev.render_nrc(cache, cam, W, H, spp=1, seed=11, spp_offset=warm + frame, accumulate=k > 0,
                                    env_is=env_is, **kw)
                frame += 1
            nr.synchronize()
            fit[kk] = ((time.perf_counter() - t0) * 1e3 / kk, None)
            dev_spp[0] = kk
            fit[kk] = (fit[kk][0], mse())
        s2 = (fit[k_small][1] - fit[k_large][1]) / (1.0 / k_small - 1.0 / k_large)
        b2 = fit[k_large][1] - s2 / k_large
        nrc_ms = fit[k_large][0]
        nrc_time = warm_s + nrc_ms * 1e-3 * s2 / (eps ** 2 - b2) if b2 < eps ** 2 else None
        return {"pt_s": pt_time, "nrc_s": nrc_time, "speedup": (pt_time / nrc_time) if nrc_time else None,
                "pt_frame_ms": pt_ms, "pt_var_per_frame": pt_s2, "nrc_frame_ms": nrc_ms, "nrc_var_per_frame": s2,
                "nrc_bias2": b2, "nrc_warmup_s": warm_s, "nrc_queries_per_frame": st["queries"],
                "nrc_trained_per_frame": st["trained"]}

    out = {"eps": eps, **legs(False)}
    out["with_sky_sampling"] = legs(True)
    out.update({"reference_spp": ref_spp, "reference_s": ref_s,
                "workload": "1920x1080, 1 spp per frame, heightfield + spheres scene, maxdepth 10; time to RMSE "
                            "extrapolated from measured per-frame variance and cache bias; with_sky_sampling: "
                            "both legs with nr_params.env_is (same expected image)"})
    return out


BF16_PEAK_TFLOPS = 2500.0
NRC_FWD_FLOP = 2 * (44 * 64 + 4 * 64 * 64 + 64 * 3)     # 38,784 per query
NRC_TRAIN_FLOP = 3 * NRC_FWD_FLOP                         # forward, input gradient, weight gradient


def nrc_kernels(nr, n_query=1920 * 1080, reps=20):
    """Radiance-cache MLP kernels on device-resident synthetic queries: one inference launch over
    a 1080p frame's worth of queries and one training step of 65,536 samples, against the BF16
    MFMA peak (algorithmic FLOPs: 2 * (44*64 + 4*64*64 + 64*3) per query, 3x that per training
    sample)."""
    cache = nr.Nrc(seed=3)
    infer_ms, train_ms = cache.bench(n_query, reps)
    inf_tf = n_query * NRC_FWD_FLOP / (infer_ms * 1e-3) / 1e12
    tr_tf = 65536 * NRC_TRAIN_FLOP / (train_ms * 1e-3) / 1e12
    return {"infer_queries": n_query, "infer_ms": infer_ms, "infer_tflops": inf_tf,
            "infer_frac_bf16_peak": inf_tf / BF16_PEAK_TFLOPS, "train_samples": 65536, "train_step_ms": train_ms,
            "train_tflops": tr_tf, "train_frac_bf16_peak": tr_tf / BF16_PEAK_TFLOPS}


def large_scene(nr, scenes, spp=256):
    """The large textured scene at 3840x2160 (9,188,616 triangles, 16 RGBA8 4096^2 textures, a
    2048^2 float texture, 10 % GGX dielectric spheres), maxdepth 12, RR from vertex 3: one frame at
    `spp` samples per pixel; the adjoint of the 2048^2 texture (12.6 M parameters, recompute-only)
    over a 1920x1080 x 4 spp image; the radiance-cache MLP over a 3840x2160 frame of queries."""
    t0 = time.perf_counter()
    sc = scenes.large_textured()
    gen_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    dev = nr.Scene(sc)
    prep_s = time.perf_counter() - t0
    W, H = sc["width"], sc["height"]
    kw = dict(maxdepth=12, rrdepth=3)
    dev.render(sc["camera"], W, H, 4, seed=1, fetch=False, **kw)
    nr.synchronize()
    t0 = time.perf_counter()
    dev.render(sc["camera"], W, H, spp, seed=2, fetch=False, **kw)
    nr.synchronize()
    frame_s = time.perf_counter() - t0
    aw, ah, aspp = 1920, 1080, 4
    cam = dict(sc["camera"], aspect=aw / ah)
    dldi = np.random.Generator(np.random.PCG64(1)).normal(size=(ah, aw, 3)).astype(np.float32)
    M = len(sc["mats"])
    dev.adjoint(cam, aw, ah, 1, M, seed=3, dl_dimg=dldi, tex_id=16, tex_shape=(2048, 2048), **kw)
    nr.synchronize()
    t0 = time.perf_counter()
    dev.adjoint(cam, aw, ah, aspp, M, seed=3, dl_dimg=dldi, tex_id=16, tex_shape=(2048, 2048), **kw)
    nr.synchronize()
    adj_s = time.perf_counter() - t0
    cache = nr.Nrc(seed=3)
    infer_ms, _ = cache.bench(W * H, 10)
    return {"triangles": int(len(sc["p"])), "generate_s": gen_s, "prepare_s": prep_s,
            "bvh": dev.info(), "frame": f"{W}x{H}x{spp}spp", "frame_s": frame_s,
            "msamples_per_s": W * H * spp / frame_s / 1e6,
            "adjoint_texture_2048_ms": adj_s * 1e3, "adjoint_samples": aw * ah * aspp,
            "adjoint_msamples_per_s": aw * ah * aspp / adj_s / 1e6, "nrc_infer_4k_ms": infer_ms}


def reference_scene(nr, scenes, with_cpu, reps=5):
    """The reference's own CPU-sized case: the 36-triangle box room at 320x240 x 16 spp with the
    reference's defaults (Russian roulette at every vertex, prr 0.8, no depth limit), timed as a
    wavefront render and as a path-kernel render (one lane per path, the same image), and, on the
    host, the oracle's restatement of the reference estimator on the same frame (all threads)."""
    sc = scenes.box_room()
    W, H, spp = sc["width"], sc["height"], 16
    dev = nr.Scene(sc)
    kw = dict(maxdepth=0, rrdepth=0, fetch=False)
    out = {"workload": f"box_room (36 triangles) {W}x{H}x{spp}spp, reference defaults (RR at every vertex, "
                       f"prr 0.8, no depth limit)", "samples": W * H * spp}
    for name, pk in (("wavefront", False), ("path_kernel", True)):
        dev.render(sc["camera"], W, H, spp, seed=1, path_kernel=pk, **kw)
        nr.synchronize()
        t0 = time.perf_counter()
        for k in range(reps):
            dev.render(sc["camera"], W, H, spp, seed=2 + k, path_kernel=pk, **kw)
        nr.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        out[f"{name}_ms"] = ms
        out[f"{name}_msamples_per_s"] = W * H * spp / ms / 1e3
    dev.close()
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po
        from nuray_amd.scenes import camera_array
        threads = len(os.sched_getaffinity(0))
        o = po.Scene(sc)
        t0 = time.perf_counter()
        o.render(camera_array(sc["camera"]), W, H, spp, seed=2, nthreads=threads)
        dt = time.perf_counter() - t0
        out["cpu_ms"] = dt * 1e3
        out["cpu_msamples_per_s"] = W * H * spp / dt / 1e6
        out["cpu_threads"] = threads
    return out


def render_modes(nr, scenes, dev, sc, reps=3):
    """Wavefront against one lane per path (nr_params.path_kernel 0 / 1, the same image) on frames
    either side of the automatic mode's samples x triangles threshold (2e12), maxdepth 10: the
    114 k-triangle terrain at 256^2 x 32 and 1280x720 x 16, and the benchmark scene (684 k) at
    256^2 x 32 and 640x360 x 16."""
    mid = scenes.terrain_spheres(n_cells=192, n_spheres=500)
    dmid = nr.Scene(mid)
    rows = []
    for d, s, frames in ((dmid, mid, ((256, 256, 32), (1280, 720, 16))), (dev, sc, ((256, 256, 32), (640, 360, 16)))):
        for W, H, spp in frames:
            cam = dict(s["camera"], aspect=W / H)
            row = {"triangles": len(s["p"]), "frame": f"{W}x{H}x{spp}", "samples_x_triangles": W * H * spp * len(s["p"])}
            for name, pk in (("wavefront", False), ("path_kernel", True)):
                d.render(cam, W, H, spp, seed=1, maxdepth=10, rrdepth=3, fetch=False, path_kernel=pk)
                nr.synchronize()
                t0 = time.perf_counter()
                for k in range(reps):
                    d.render(cam, W, H, spp, seed=2 + k, maxdepth=10, rrdepth=3, fetch=False, path_kernel=pk)
                nr.synchronize()
                row[f"{name}_ms"] = (time.perf_counter() - t0) * 1e3 / reps
            rows.append(row)
    dmid.close()
    return rows


def env_sampling(nr, dev, sc, spp):
    """The same 720p frame with and without sky sampling (nr_params.env_is): time per frame, mean
    per-pixel variance of one sample, and the time-to-equal-error gain var_off t_off / (var_on t_on)."""
    W, H = sc["width"], sc["height"]
    kw = dict(spp=spp, seed=11, maxdepth=10, rrdepth=3, prr=0.8)
    res = {}
    for env in (False, True):
        dev.render(sc["camera"], W, H, fetch=False, env_is=env, **kw)   # buffers sized for this mode
        nr.synchronize()
        t0 = time.perf_counter()
        dev.render(sc["camera"], W, H, fetch=False, env_is=env, **kw)
        nr.synchronize()
        dt = time.perf_counter() - t0
        r = dev.fetch(W, H)
        m = r["sum"] / spp
        var = float(np.mean(np.maximum(r["sum2"] / spp - m * m, 0) * spp / (spp - 1)))
        res[env] = (dt, var, float(m.mean()))
    (t0_, v0, m0), (t1, v1, m1) = res[False], res[True]
    return {"frame": f"{W}x{H}x{spp}spp", "off_ms": t0_ * 1e3, "on_ms": t1 * 1e3,
            "off_msamples_per_s": W * H * spp / t0_ / 1e6, "on_msamples_per_s": W * H * spp / t1 / 1e6,
            "var_off": v0, "var_on": v1, "mean_off": m0, "mean_on": m1,
            "variance_ratio": v0 / v1, "equal_error_speedup": v0 * t0_ / (v1 * t1)}


def profile_figures(name):
    """key -> value lines of a committed rocprofv3 summary under profiles/ (counter runs cannot be
    taken inside the timed run: a profiler slows the kernels it counts)."""
    path = os.path.join(ROOT, "profiles", name)
    out = {}
    if os.path.exists(path):
        for line in open(path):
            parts = line.split()
            if len(parts) >= 2 and not line.startswith("#"):
                try:
                    out[parts[0]] = float(parts[1])
                except ValueError:
                    out[parts[0]] = " ".join(parts[1:])
    return out


def traversal_roofline(cnt, closest_ms, launches, steps):
    """Roofline of the dominant kernel, the closest-hit traversal k_trace<closest>.

    Its counters show it bound by VALU issue, not by HBM: the exactness contract fixes the work
    (every box the reference's slab test accepts is visited), and the node and triangle bytes are
    mostly served by the L2s and the Infinity Cache.  So the roofline is VALU issue against 39.3 T
    lane-operations/s, and `achieved` is algorithmic: the reference's FP32 operations per box test
    and per triangle test (OPS_NODE, OPS_TRI) times the box and triangle tests of this run's own
    counter pass, over this run's launch time.  Beside it, `lane_slots`: the lane-operations the
    kernel actually issues (VALU instructions x active lanes, bookkeeping included) from the
    committed counter run profiles/trace_closest_720p_lane_ops.txt, over the same rays and time.
    HBM: the measured traffic per launch (FETCH_SIZE / WRITE_SIZE counters calibrated on a copy
    of known size, profiles/trace_closest_720p_traffic.txt) over this run's launch time.  The
    no-reuse byte count (128 B per node fetch, 48 B per triangle test and per ray) is reported as
    bytes_no_reuse_per_launch: it counts cache hits as traffic, so it is not a fraction of HBM peak."""
    avg_ms = closest_ms / max(launches, 1)
    n_launch = max(cnt["closest_launches"], 1)
    rays_launch = cnt["rays"] / n_launch
    alg_ops = (OPS_NODE * cnt["node_fetch"] + OPS_TRI * cnt["tri_tests"]) / n_launch
    achieved = alg_ops / (avg_ms * 1e-3) / 1e12 if avg_ms > 0 else None
    lane = profile_figures("trace_closest_720p_lane_ops.txt")
    lane_ops_ray = lane.get("lane_ops_per_ray")
    issued = lane_ops_ray * rays_launch / (avg_ms * 1e-3) / 1e12 if lane_ops_ray and avg_ms > 0 else None
    hbm = profile_figures("trace_closest_720p_traffic.txt")
    traffic = hbm.get("traffic_bytes_per_launch")
    hbm_gbs = traffic / (avg_ms * 1e-3) / 1e9 if traffic and avg_ms > 0 else None
    no_reuse = (B_NODE * cnt["node_fetch"] + B_TRI * cnt["tri_tests"] + B_RAY * cnt["rays"]) / n_launch
    return {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_TLANEOPS, "unit": "T lane-ops/s",
            "frac": achieved / VALU_PEAK_TLANEOPS if achieved else None, "traffic": traffic,
            "kernel": "k_trace<closest>", "avg_launch_ms": avg_ms, "launches_per_frame": launches / steps,
            "launch_times_from": "one untimed frame with HIP events and serial launches",
            "rays_per_launch": rays_launch, "alg_ops_per_launch": alg_ops,
            "alg_ops_per_node_fetch": OPS_NODE, "alg_ops_per_tri_test": OPS_TRI,
            "lane_slots": {"achieved": issued, "frac": issued / VALU_PEAK_TLANEOPS if issued else None,
                           "lane_ops_per_ray": lane_ops_ray, "valu_busy_in_counter_run": lane.get("valu_busy"),
                           "active_lanes_per_valu_inst": lane.get("active_lanes_per_inst"),
                           "source": lane.get("source")},
            "hbm": {"achieved": hbm_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": hbm_gbs / HBM_PEAK_GBS if hbm_gbs else None, "traffic": traffic,
                    "traffic_source": hbm.get("source")},
            "bytes_no_reuse_per_launch": no_reuse,
            "no_reuse_bytes_per_s_over_hbm_peak": no_reuse / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if avg_ms > 0 else None,
            "rays_per_frame": cnt["rays"], "node_fetch_per_ray": cnt["node_fetch"] / max(cnt["rays"], 1),
            "tri_tests_per_ray": cnt["tri_tests"] / max(cnt["rays"], 1),
            "rays_retried_with_full_stack": cnt["retried"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-nrc", action="store_true", help="skip the time-to-RMSE measurement (about 40 s)")
    ap.add_argument("--frames-only", action="store_true",
                    help="only the ptnee frames (no adjoint, cache or CPU legs), e.g. under a profiler")
    ap.add_argument("--serial-frames", action="store_true",
                    help="run the warm-up and timed frames with HIP-event timing on, so that their launches run one "
                         "after another as in the kernel-time frame (for kernel-trace profiles, whose per-dispatch "
                         "durations then match the roofline's launch time); the default overlaps each bounce's "
                         "shadow launch with the next closest-hit launch")
    ap.add_argument("--trace-order", type=int, default=3,
                    help="order of the wavefront trace launches: 0 queue, 1 octant per 1024 rays, 2 64 direction "
                         "bins per 4096 rays, 3 (the library's default) shadow rays alone by 64 bins per 256 rays, "
                         "4 the same per 1024 rays (same results)")
    ap.add_argument("--large", action="store_true",
                    help="only measure the large textured scene (3840x2160 x 256 spp, texel adjoint) and exit")
    ap.add_argument("--mlp-only", action="store_true",
                    help="only time the radiance-cache MLP kernels (inference + training step) and exit")
    ap.add_argument("--copy-only", action="store_true",
                    help="only run the plain 1 GiB device copy (HBM counter calibration) and exit")
    ap.add_argument("--atomics-only", action="store_true",
                    help="only measure device-scope atomic adds on counters sharing a 128-byte line or not, "
                         "and exit")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only time the inverse-rendering iteration (and the reference box room) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "nuray_amd runs on one GPU: a frame shards by pixels x samples, but "
                                   "multi-GPU runs are not implemented", "n_gpus": args.gpus}))
        sys.exit(2)

    import nuray_amd as nr
    from nuray_amd import scenes

    nr.init(0)
    copy_gbs = nr.bench_copy(1 << 30, 5)
    if args.copy_only:
        print(json.dumps({"copy_bytes": 1 << 30, "reps": 5, "copy_gbs": copy_gbs}), flush=True)
        return
    if args.atomics_only:
        # (counters, ints apart): one counter; two in one 128-byte line; two, eight and 32 on lines of
        # their own
        cases = [(1, 1), (2, 1), (2, 32), (8, 32), (32, 32)]
        print(json.dumps({"atomic_adds_mops": [{"counters": c, "bytes_apart": 4 * s,
                                                 "mops": nr.bench_atomics(c, s)} for c, s in cases]}), flush=True)
        return
    if args.mlp_only:
        print(json.dumps(nrc_kernels(nr)), flush=True)
        return
    if args.large:
        print(json.dumps(large_scene(nr, scenes)), flush=True)
        return
    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_iteration(nr, scenes),
                          "reference_scene": reference_scene(nr, scenes, False)}), flush=True)
        return
    sc = scenes.terrain_spheres()
    W, H, spp = sc["width"], sc["height"], args.spp
    t0 = time.perf_counter()
    dev = nr.Scene(sc)
    prep_s = time.perf_counter() - t0
    dev.set_trace_order(args.trace_order)
    info = dev.info()
    rp = dict(spp=spp, seed=7, maxdepth=10, rrdepth=3, prr=0.8, fetch=False)
    print(f"# scene {sc['name']}: {len(sc['p'])} triangles, BVH {info['n_nodes']} nodes depth {info['depth']}, "
          f"build {info['build_ms']:.0f} ms, prepare {prep_s:.2f} s", file=sys.stderr, flush=True)

    dev.set_profiling(1 if args.serial_frames else 0)
    for _ in range(args.warmup):
        dev.render(sc["camera"], W, H, **rp)
    # counters pass (untimed): node fetches / triangle tests / rays of the closest-hit kernel
    dev.set_profiling(2)
    dev.render(sc["camera"], W, H, **rp)
    cnt = dev.stats()
    # kernel-time pass (untimed): with HIP events on, the launches run one after another, so each
    # kernel's time is its own; the timed frames below overlap a bounce's shadow launch with the
    # next bounce's closest-hit launch.
    dev.set_profiling(1)
    dev.render(sc["camera"], W, H, **rp)
    kt = dev.stats()
    closest_ms, shade_ms, shadow_ms = kt["closest_ms"], kt["shade_ms"], kt["shadow_ms"]
    dev.set_profiling(1 if args.serial_frames else 0)
    nr.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dev.render(sc["camera"], W, H, **rp)
    nr.synchronize()
    elapsed = time.perf_counter() - t0
    frame_mean = dev.fetch(W, H)["sum"] / spp   # the timed frames' image (same seed every frame)
    samples = W * H * spp * args.steps
    value = samples / elapsed / 1e6
    print(f"# timed {args.steps} frames in {elapsed:.3f} s; kernel times of one serial frame: closest-hit "
          f"{closest_ms:.1f} ms, shade {shade_ms:.1f} ms, shadow {shadow_ms:.1f} ms; "
          f"{kt['iterations']:.0f} iterations", file=sys.stderr, flush=True)

    roofline = traversal_roofline(cnt, closest_ms, kt["closest_launches"], 1)
    if args.frames_only:
        args.no_cpu = args.no_nrc = True
    adj = None if args.frames_only else adjoint_iteration(nr, scenes)
    if adj:
        print(f"# adjoint iteration {adj['ms_per_iter']:.2f} ms (render {adj['render_ms']:.2f} ms, "
              f"replay {adj['adjoint_ms']:.2f} ms; wavefront render {adj['wavefront_render_ms']:.2f} ms)",
              file=sys.stderr, flush=True)
    env = None if args.frames_only else env_sampling(nr, dev, sc, spp)
    if env:
        print(f"# sky sampling: {env['on_ms']:.1f} ms vs {env['off_ms']:.1f} ms per frame, variance "
              f"{env['variance_ratio']:.2f}x lower, {env['equal_error_speedup']:.2f}x faster to equal error",
              file=sys.stderr, flush=True)
    ref_scene = None if args.frames_only else reference_scene(nr, scenes, not args.no_cpu)
    modes = None if args.frames_only else render_modes(nr, scenes, dev, sc)
    if modes:
        print("# render modes (wavefront / one lane per path, ms): " + ", ".join(
            f"{m['triangles']} tris {m['frame']} {m['wavefront_ms']:.2f} / {m['path_kernel_ms']:.2f}" for m in modes),
            file=sys.stderr, flush=True)
    if ref_scene:
        print(f"# reference box room 320x240x16spp: wavefront {ref_scene['wavefront_ms']:.2f} ms, path kernel "
              f"{ref_scene['path_kernel_ms']:.2f} ms"
              + (f", CPU oracle {ref_scene['cpu_ms']:.0f} ms" if "cpu_ms" in ref_scene else ""),
              file=sys.stderr, flush=True)
    nrc = None if args.no_nrc else nrc_time_to_rmse(nr, scenes)
    mlp = None if args.frames_only else nrc_kernels(nr)
    if mlp:
        print(f"# cache MLP: inference {mlp['infer_ms']:.3f} ms / 1080p frame ({mlp['infer_tflops']:.0f} TFLOP/s), "
              f"training step {mlp['train_step_ms']:.3f} ms ({mlp['train_tflops']:.0f} TFLOP/s)", file=sys.stderr,
              flush=True)
    if nrc:
        for lab, r in (("", nrc), (" with sky sampling", nrc["with_sky_sampling"])):
            print(f"# time to RMSE 1e-2 at 1080p{lab}: PT {r['pt_s']:.1f} s, radiance cache "
                  f"{r['nrc_s'] if r['nrc_s'] is None else round(r['nrc_s'], 1)} s", file=sys.stderr, flush=True)
    cpu = None if args.no_cpu else cpu_baseline(sc, spp, frame_mean)
    out = {
        "metric": "Msamples/s at 720p", "value": value, "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"ptnee {W}x{H}x{spp}spp maxdepth10 rrdepth3 prr0.8, {sc['name']} "
                               f"({len(sc['p'])} triangles) + 1024x512 sky", "global_batch": W * H * spp,
                   "seq_len": None, "parallelism": "single-gpu", "bvh_build_ms": info["build_ms"],
                   "trace_order": args.trace_order,
                   "launches": "serial" if args.serial_frames else "shadow beside the next closest hit"},
        "roofline": roofline, "cpu_baseline": cpu, "hbm_copy_gbs": copy_gbs,
        "kernel_ms_per_frame": {"closest_hit": closest_ms, "shade": shade_ms, "shadow": shadow_ms,
                                "from": "one untimed frame with serial launches"},
        "shadow_rays": {"rays_per_frame": cnt["shadow_rays"],
                        "node_fetch_per_ray": cnt["shadow_node_fetch"] / max(cnt["shadow_rays"], 1),
                        "tri_tests_per_ray": cnt["shadow_tri_tests"] / max(cnt["shadow_rays"], 1)},
        "adjoint": adj, "nrc": nrc, "nrc_mlp": mlp, "env_sampling": env, "reference_scene": ref_scene,
        "render_modes": modes,
    }
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
        out["mre_fp64"] = cpu["mre_fp64"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
