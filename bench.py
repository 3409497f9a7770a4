"""Benchmark: NCC search points/s at chip 32, search radius 25, 64x subpixel refinement.

Workload: a 10000x10000 seeded band-limited texture (Gaussian blur sigma 1.5 px), the secondary
warped by a smooth field (sum of 8 Gaussians, peak 20 px, bilinear) plus N(0, 2^2) noise; both
prepared on the device by the default pipeline (high-pass width 5, UInt8 quantisation); grid spacing 24 px with the margin of the
largest chip level (128), i.e. 410 x 410 = 168 100 search points, each correlated with a 32 x 32
chip over a 50 x 50 shift surface and refined 64x. One step = one tracking pass over all points
with images and points already resident in HBM.

Prints one JSON line (see the driver contract): value = searched points per second.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "autorift.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

INT8_PEAK_TOPS = 5000.0          # MI355X dense int8 MFMA, TOP/s
FLOP_PER_POINT = 2 * 50 * 50 * 32 * 32   # useful multiply-adds of the direct contraction, x2
# rocprofv3 FETCH_SIZE/WRITE_SIZE summary of the default workload, scaled by a 1 GiB copy
TRAFFIC_PROFILE = os.path.join("profiles", "track_u8_kernel_10000sq_c32r25.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def synthetic_pair(n, seed=3, fields=False):
    """Seeded texture and a smoothly warped, noisy copy (float32); with `fields`, also the
    displacement (fx, fy) in pixels: secondary(y, x) = texture(y - fy, x - fx)."""
    from scipy import ndimage
    rng = np.random.default_rng(seed)
    tex = rng.standard_normal((n, n), dtype=np.float32)
    tex = ndimage.gaussian_filter(tex, 1.5, truncate=4.0, mode="mirror")
    tex -= tex.min()
    tex *= 255.0 / tex.max()
    # smooth displacement field: 8 Gaussians, scaled so max |d| = 20 px
    frng = np.random.default_rng(seed + 1)
    m = 256                       # evaluate coarsely, then upsample (field is smooth)
    yy, xx = np.meshgrid(np.linspace(0, n, m), np.linspace(0, n, m), indexing="ij")
    fx = np.zeros((m, m))
    fy = np.zeros((m, m))
    for _ in range(8):
        cy, cx = frng.uniform(0, n, 2)
        s = frng.uniform(500, 2000)
        ax, ay = frng.normal(0, 1, 2)
        g = np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * s * s))
        fx += ax * g
        fy += ay * g
    scale = 20.0 / max(np.abs(fx).max(), np.abs(fy).max())
    fx = ndimage.zoom(fx * scale, n / m, order=1).astype(np.float32)
    fy = ndimage.zoom(fy * scale, n / m, order=1).astype(np.float32)
    sec = np.empty_like(tex)
    step = 1000
    for r0 in range(0, n, step):
        r1 = min(n, r0 + step)
        I, J = np.meshgrid(np.arange(r0, r1, dtype=np.float32), np.arange(n, dtype=np.float32),
                           indexing="ij")
        coords = np.stack([I - fy[r0:r1], J - fx[r0:r1]])
        sec[r0:r1] = ndimage.map_coordinates(tex, coords, order=1, mode="nearest")
    sec += np.random.default_rng(seed + 2).normal(0, 2, sec.shape).astype(np.float32)
    if fields:
        return tex, sec, fx, fy
    return tex, sec


def cfg1_pair(seed, n=512, shift=(3.30, -1.70)):
    """512^2 blurred white noise scaled to 0..255; the secondary is the reference Fourier-shifted
    by (dx, dy) plus N(0, 2^2) noise (Float32)."""
    from scipy import ndimage
    rng = np.random.default_rng(seed)
    a = ndimage.gaussian_filter(rng.standard_normal((n, n)), 1.5, truncate=4.0, mode="wrap")
    a = (a - a.min()) * (255.0 / (a.max() - a.min()))
    ky = np.fft.fftfreq(n)[:, None]
    kx = np.fft.fftfreq(n)[None, :]
    dx, dy = shift
    b = np.real(np.fft.ifft2(np.fft.fft2(a) * np.exp(-2j * np.pi * (kx * dx + ky * dy))))
    b = b + rng.normal(0, 2, b.shape)
    return a.astype(np.float32), b.astype(np.float32)


def pair_latency(autorift, npairs, warm=2):
    """Default pipeline on 512^2 pairs (high-pass, UInt8, multi-chip search over the 784-point
    grid 40:16:472), one context reused across pairs as init/reinit!/autorift! do. The images are
    held column-major, as a Julia caller holds them (the ABI's layout: no host transposition)."""
    xs = np.arange(40.0, 473.0, 16.0)
    grid = autorift.gridpoints_xy(xs, xs, chip_size=32, search_radius=25)
    pairs = [tuple(np.asfortranarray(x) for x in cfg1_pair(100 + k)) for k in range(npairs + warm)]
    ctx = autorift.Context(autorift.params())
    lat = []
    meas = []
    for k, (a, b) in enumerate(pairs):
        t = time.perf_counter()
        ctx.set_images(a, b)
        r = ctx.run_multichip(grid)
        el = time.perf_counter() - t
        if k >= warm:
            lat.append(el * 1000)
            meas.append((np.nanmedian(r.dx), np.nanmedian(r.dy), r.nmeasured()))
    ctx.close()
    lat = np.array(lat)
    return {"pairs": npairs, "grid_points": len(grid), "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "pairs_per_s": float(1000 / lat.mean()),
            "median_dx": float(np.median([m[0] for m in meas])),
            "median_dy": float(np.median([m[1] for m in meas])),
            "expected_dx_dy": [-3.30, 1.70]}


def serve_pool(count, seed=6, lo=512, hi=10000, sides=()):
    """Pairs for the serving sweep: `count` sides log-uniform in [lo, hi], then the given `sides`;
    blurred noise (sigma 1.5) scaled to 0..255, the secondary translated by a random subpixel
    shift (|d| <= 15 px, bilinear) plus N(0, 2^2) noise; Float32."""
    from scipy import ndimage
    rng = np.random.default_rng(seed)
    out = []
    todo = [None] * count + list(sides)
    for side in todo:
        n = side or int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        a = ndimage.gaussian_filter(rng.standard_normal((n, n), dtype=np.float32), 1.5,
                                    truncate=4.0, mode="mirror")
        a -= a.min()
        a *= 255.0 / a.max()
        ang, mag = rng.uniform(0, 2 * np.pi), rng.uniform(0, 15)
        d = (mag * np.sin(ang), mag * np.cos(ang))            # (dy, dx)
        b = ndimage.shift(a, d, order=1, mode="nearest")
        b += rng.normal(0, 2, b.shape).astype(np.float32)
        out.append((np.asfortranarray(a), np.asfortranarray(b), d))   # the ABI's layout
    return out


def serving_sweep(autorift, rates, nreq_max, slots, pool_size, seed=6):
    """Poisson arrivals of pairs from the pool (pool_size random sides and one 10000^2 pair) at
    each rate into one Server; 3 s of arrivals per rate, between 10 and nreq_max requests; per
    rate: p50 / p99 latency (submission to result), pairs/s, and the server's peak
    concurrency."""
    import threading
    t0 = time.time()
    pool = serve_pool(pool_size, seed, sides=(10000,))
    log(f"serving pool: sides {[a.shape[0] for a, _, _ in pool]} in {time.time() - t0:.1f}s")
    p = autorift.params()
    rng = np.random.default_rng(seed + 1)
    res = {"slots": slots, "pool_sides": [int(a.shape[0]) for a, _, _ in pool], "rates": []}
    with autorift.Server(p, slots=slots) as srv:
        for a, b, _ in pool[:2]:          # warm the slots' buffers and kernels
            srv.result(srv.submit(a, b))
        for rate in rates:
            nreq = int(min(nreq_max, max(10, round(3 * rate))))
            order = rng.integers(0, len(pool), nreq)
            gaps = rng.exponential(1.0 / rate, nreq)
            tickets = [None] * nreq
            ready = threading.Semaphore(0)

            def submit():
                t = time.perf_counter()
                for k in range(nreq):
                    t += gaps[k]
                    d = t - time.perf_counter()
                    if d > 0:
                        time.sleep(d)
                    a, b, _ = pool[order[k]]
                    tickets[k] = srv.submit(a, b)
                    ready.release()

            th = threading.Thread(target=submit)
            start = time.perf_counter()
            th.start()
            lat, errs = [], []
            for k in range(nreq):
                ready.acquire()
                r, tm = srv.result(tickets[k])
                lat.append(tm["latency_ms"])
                dy, dx = pool[order[k]][2]
                ok = np.isfinite(r.dx)
                if ok.any():   # outputs are secondary -> reference offsets
                    errs.append(float(np.hypot(np.median(r.dx[ok]) + dx,
                                               np.median(r.dy[ok]) + dy)))
            th.join()
            wall = time.perf_counter() - start
            st = srv.stats()
            lat = np.array(lat)
            row = {"rate": rate, "requests": nreq, "p50_ms": float(np.percentile(lat, 50)),
                   "p99_ms": float(np.percentile(lat, 99)), "pairs_per_s": nreq / wall,
                   "peak_running": st["peak_running"],
                   "median_shift_error_px": float(np.median(errs)) if errs else None}
            res["rates"].append(row)
            log(f"serve rate {rate}/s: p50 {row['p50_ms']:.1f} ms p99 {row['p99_ms']:.1f} ms, "
                f"{row['pairs_per_s']:.1f} pairs/s, peak running {st['peak_running']}")
        res["peak_bytes_held"] = srv.stats()["peak_bytes_held"]
    return res


def served_mixed_vs_plain(autorift, count=8, slots=4, seed=7, reps=3):
    """Float32 pairs (quantize = :none), default multi-chip search: the plain form (exact Float64
    contraction, one pair at a time on one context) against the served mixed-precision form
    (BF16 contraction, pairs submitted together to a Server with `slots` slots): pairs/s of each
    (the best of `reps` timed passes over the pool), and how far the served results are from the
    plain ones."""
    pool = serve_pool(count, seed, lo=1024, hi=4096)
    log(f"served-vs-plain pool: sides {[a.shape[0] for a, _, _ in pool]}")
    plain = []
    # Each form is warmed up on the whole pool first (untimed), so both are timed at steady
    # state: the plain context's buffers have grown to the largest pair, the server's page-locked
    # staging pool holds a buffer for every pair in flight.
    with autorift.Context(autorift.params(quantize="none")) as ctx:
        for a, b, _ in pool:                           # warm-up
            ctx.set_images(a, b)
            ctx.run_multichip(autorift.api._build_grid(a.shape, ctx.params))
        plain_s = float("inf")
        for _ in range(reps):
            t = time.perf_counter()
            set_k = 0.0
            plain = []
            for a, b, _ in pool:
                t0 = time.perf_counter()
                ctx.set_images(a, b)                   # upload and preparation of the pair
                set_k += time.perf_counter() - t0
                plain.append(ctx.run_multichip(autorift.api._build_grid(a.shape, ctx.params)))
            el = time.perf_counter() - t
            if el < plain_s:
                plain_s, set_s = el, set_k
    with autorift.Server(autorift.params(quantize="none", precision="bf16"), slots=slots) as srv:
        for k in [srv.submit(a, b) for a, b, _ in pool]:   # warm-up
            srv.result(k)
        served_s = float("inf")
        for _ in range(reps):
            t = time.perf_counter()
            tickets = [srv.submit(a, b) for a, b, _ in pool]
            served = [srv.result(k)[0] for k in tickets]
            served_s = min(served_s, time.perf_counter() - t)
    same, tot, sq = 0, 0, []
    for r, e in zip(served, plain):
        ok = np.isfinite(e.dx) & np.isfinite(r.dx)
        tot += int(ok.sum())
        same += int(np.count_nonzero((r.dx[ok] == e.dx[ok]) & (r.dy[ok] == e.dy[ok])))
        sq.append((r.dx[ok] - e.dx[ok]) ** 2 + (r.dy[ok] - e.dy[ok]) ** 2)
    res = {"pairs": count, "sides": [int(a.shape[0]) for a, _, _ in pool], "slots": slots,
           "plain_fp32_pairs_per_s": count / plain_s, "served_bf16_pairs_per_s": count / served_s,
           "speedup": plain_s / served_s, "timed_passes": reps,
           # share of the plain form's time in set_images (host-to-device upload of the two
           # Float32 images and their preparation): work the served form repeats for every pair
           "plain_set_images_share": set_s / plain_s,
           "identical_fraction": same / max(tot, 1),
           "rms_vs_plain_px": float(np.sqrt(np.mean(np.concatenate(sq)))) if sq else None}
    log(f"served BF16 {res['served_bf16_pairs_per_s']:.1f} pairs/s vs plain Float32 "
        f"{res['plain_fp32_pairs_per_s']:.2f} pairs/s ({res['speedup']:.1f}x; plain time in "
        f"upload and preparation {100 * res['plain_set_images_share']:.0f} %); identical "
        f"{res['identical_fraction']:.4f}, RMS {res['rms_vs_plain_px']:.4f} px")
    return res


def cfg2_block(autorift, tex, sec, fx, fy, reps=2):
    """The production configuration as written: the multi-chip search over chip sizes 16, 32, 64
    and 128 (search radius 25, grid spacing 24) on the 10000^2 pair, for each quantiser and
    contraction: preparation (upload of the two Float32 images included) and search times (best
    of `reps`), grid points per second of the search, points finished per chip-size level, and
    the error against the known displacement field."""
    out = {}
    # held column-major, as the ABI takes them (a Julia caller's layout: no host transposition)
    tex, sec = np.asfortranarray(tex), np.asfortranarray(sec)
    forms = (("uint8", dict(quantize="uint8")),
             ("float32_exact", dict(quantize="none")),
             ("float32_bf16", dict(quantize="none", precision="bf16")),
             ("float32_fft", dict(quantize="none", precision="fft")))
    for name, kw in forms:
        p = autorift.params(chip_size=16, chip_size_max=128, grid_spacing=24, search_radius=25,
                            **kw)
        grid = autorift.api._build_grid(tex.shape, p)
        prep_s = search_s = float("inf")
        with autorift.Context(p) as ctx:
            for _ in range(reps):
                t0 = time.perf_counter()
                ctx.set_images(tex, sec)
                t1 = time.perf_counter()
                r = ctx.run_multichip(grid)
                t2 = time.perf_counter()
                prep_s, search_s = min(prep_s, t1 - t0), min(search_s, t2 - t1)
        ok = np.isfinite(r.dx)
        iy = np.clip(np.floor(grid.y).astype(int) - 1, 0, tex.shape[0] - 1)
        ix = np.clip(np.floor(grid.x).astype(int) - 1, 0, tex.shape[1] - 1)
        err = np.hypot(r.dx + fx[iy, ix], r.dy + fy[iy, ix])[ok]
        levels = {str(int(c)): int((r.chip_size == c).sum()) for c in np.unique(r.chip_size)}
        out[name] = {"grid_points": int(grid.x.size), "prepare_ms": prep_s * 1000,
                     "search_ms": search_s * 1000,
                     "grid_points_per_s": grid.x.size / search_s,
                     "measured": float(ok.mean()), "points_per_chip_size": levels,
                     "median_error_px": float(np.median(err)),
                     "p95_error_px": float(np.percentile(err, 95))}
        log(f"cfg2 {name}: prepare {prep_s * 1000:.1f} ms, search {search_s * 1000:.1f} ms "
            f"({grid.x.size / search_s / 1e6:.2f} M grid points/s), levels {levels}, median "
            f"error {np.median(err):.3f} px")
    return out


def speckle_pair(n, shift, seed):
    """Circular Gaussian speckle smoothed by two 2x2 box passes (wrap-around), and a copy rolled
    by `shift` plus a little complex noise: a synthetic single-look-complex pair."""
    rng = np.random.default_rng(seed)
    z = (rng.standard_normal((n, n), np.float32) +
         1j * rng.standard_normal((n, n), np.float32)).astype(np.complex64)
    for _ in range(2):
        z = (z + np.roll(z, 1, 0) + np.roll(z, 1, 1) + np.roll(np.roll(z, 1, 0), 1, 1)) / 4
    z2 = np.roll(z, shift, axis=(0, 1))
    z2 = z2 + 0.05 * (rng.standard_normal((n, n), np.float32) +
                      1j * rng.standard_normal((n, n), np.float32))
    return z.astype(np.complex64), z2.astype(np.complex64)


def coherence_block(autorift, steps, n=2048):
    """Coherence over a complex pair (Deramp, quantize = :none): chip 32, r 20, 64x refinement,
    grid spacing 16; throughput of the device pass and the oracle on a sample of the points."""
    z1, z2 = speckle_pair(n, (5, -3), seed=6)
    xs = np.arange(64.0, n - 63.0, 16.0)
    pts = autorift.gridpoints_xy(xs, xs, chip_size=32, search_radius=20)
    p = autorift.params(similarity="coherence", preprocess="deramp", quantize="none")
    with autorift.Context(p) as ctx:
        ctx.set_images(z1, z2)
        ctx.track(pts, upsampling=64)
        t = []
        for _ in range(steps):
            r = ctx.track(pts, upsampling=64)
            t.append(ctx.last_kernel()[0])
        pr, ps, pm, _ = ctx.get_pair()
    ms = float(np.median(t))
    fin = np.isfinite(r.dx)
    out = {"pair": f"{n}x{n} ComplexF32 speckle rolled by (5, -3)", "points": len(pts),
           "chip": 32, "radius": 20, "upsampling": 64, "ms": ms,
           "points_per_s": int(np.count_nonzero(r.searched)) / (ms / 1000),
           "median_dx": float(np.median(r.dx[fin])), "median_dy": float(np.median(r.dy[fin])),
           "expected_dx_dy": [3.0, -5.0]}
    try:
        import oracle
        sub = pts[::12, ::12]
        t0 = time.perf_counter()
        oracle.track(pr, ps, pm, sub.x, sub.y, sub.radius_x, sub.radius_y, sub.chip_size_x,
                     sub.chip_size_y, up=64, measure=oracle.COHERENCE, nthreads=16)
        cpu_s = time.perf_counter() - t0
        out["cpu_oracle"] = {"points": len(sub), "threads": 16,
                             "points_per_s": len(sub) / cpu_s}
    except Exception as e:   # the CPU comparison is optional
        out["cpu_oracle"] = f"unavailable: {e}"
    log(f"coherence: {out['points_per_s']:.0f} points/s over {len(pts)} points, median "
        f"({out['median_dx']:.3f}, {out['median_dy']:.3f})")
    return out


FP64_PEAK_TFLOPS = 78.6          # MI355X FP64 vector (and matrix), TFLOP/s
FP32_PEAK_TFLOPS = 157.3         # MI355X FP32 (vector and matrix), TFLOP/s
BF16_PEAK_TFLOPS = 2500.0        # MI355X dense BF16 / FP16 matrix, TFLOP/s


def mixed_block(autorift, tex, sec, pts, steps):
    """The Float32 pair (quantize = :none) over the full grid: the exact direct Float64 tracker
    (the plain Float32 drop-in, bit-exact), the single-precision FFT tracker (the reference's own
    numerator), and the BF16 / FP16 matrix-core tracker, each with the roofline of its
    arithmetic (the FFT's as direct-equivalent multiply-adds against the FP32 peak); the
    one-kernel generic Float64 tracker on every 7th grid row and column for comparison;
    agreement on the 512^2 shifted pair."""
    import os as _os
    out = {}
    flop = FLOP_PER_POINT
    for prec in ("bf16", "fp16", "fft", "exact", "generic"):
        ctx = autorift.Context(autorift.params(quantize="none",
                                               precision="exact" if prec == "generic" else prec))
        ctx.set_images(tex, sec)
        use = pts if prec != "generic" else pts[::7, ::7]
        if prec == "generic":
            _os.environ["AR_F64_DIRECT"] = "never"
        try:
            ctx.track(use, upsampling=64)                 # warm-up (box images, kernels)
            t, tt = [], []
            for _ in range(steps if prec != "generic" else 1):
                r = ctx.track(use, upsampling=64)
                box_ms, track_ms, _ = ctx.last_timing()   # box: Float64 box images (mixed path)
                t.append(ctx.last_kernel()[0] + box_ms)
                tt.append(track_ms)
        finally:
            _os.environ.pop("AR_F64_DIRECT", None)
        ms, ms_t = float(np.median(t)), float(np.median(tt))
        n_s = int(np.count_nonzero(r.searched))
        out[prec] = {"points": len(use), "ms": ms, "track_ms": ms_t,
                     "points_per_s": n_s / (ms / 1000), "path": ctx.last_kernel()[1]}
        peak = (FP64_PEAK_TFLOPS if prec in ("exact", "generic") else
                FP32_PEAK_TFLOPS if prec == "fft" else BF16_PEAK_TFLOPS)
        ach = n_s * flop / (ms_t / 1000) / 1e12
        # the tracker phase: HIP events on the pass's stream around the correlation launches
        # (for the direct Float64 tracker: the first chunk's preparation, then every
        # track_f64_kernel launch; later chunks' preparation overlaps on a second stream)
        out[prec]["roofline"] = {"bound": {FP64_PEAK_TFLOPS: "fp64", FP32_PEAK_TFLOPS: "fp32"}
                                 .get(peak, "mfma"),
                                 "achieved": ach, "peak": peak, "unit": "TFLOP/s",
                                 "frac": ach / peak, "scope": "tracker phase"}
        if prec == "fft":
            out[prec]["roofline"]["work"] = "direct-equivalent multiply-adds (5.12 MFLOP a point)"
        ctx.close()
        log(f"mixed {prec}: {out[prec]['points_per_s']:.0f} points/s over {len(use)} points "
            f"({ach:.1f} TFLOP/s)")
    out["speedup_bf16_vs_exact"] = out["bf16"]["points_per_s"] / out["exact"]["points_per_s"]
    out["speedup_fft_vs_exact"] = out["fft"]["points_per_s"] / out["exact"]["points_per_s"]
    out["speedup_exact_vs_generic"] = (out["exact"]["points_per_s"] /
                                       out["generic"]["points_per_s"])
    # agreement on the 512^2 pair shifted by (3.30, -1.70)
    a, b = cfg1_pair(1)
    xs = np.arange(40.0, 473.0, 16.0)
    g = autorift.gridpoints_xy(xs, xs, chip_size=32, search_radius=25)
    res = {}
    for prec in ("exact", "bf16", "fp16", "fft"):
        with autorift.Context(autorift.params(quantize="none", precision=prec)) as c:
            c.set_images(a, b)
            res[prec] = (c.track(g, upsampling=1), c.track(g, upsampling=64))
    e1, e64 = res["exact"]
    ok = np.isfinite(e64.dx)
    for prec in ("bf16", "fp16", "fft"):
        r1, r64 = res[prec]
        same = (r1.dx[ok] == e1.dx[ok]) & (r1.dy[ok] == e1.dy[ok])
        err = np.hypot(r64.dx[ok] - e64.dx[ok], r64.dy[ok] - e64.dy[ok])
        tru = np.hypot(r64.dx[ok] + 3.30, r64.dy[ok] - 1.70)
        out[prec]["cfg1_peak_agreement"] = float(same.mean())
        out[prec]["cfg1_subpixel_rms_vs_fp64_px"] = float(np.sqrt(np.mean(err ** 2)))
        out[prec]["cfg1_rms_vs_known_shift_px"] = float(np.sqrt(np.mean(tru ** 2)))
    out["exact"]["cfg1_rms_vs_known_shift_px"] = float(
        np.sqrt(np.mean(np.hypot(e64.dx[ok] + 3.30, e64.dy[ok] - 1.70) ** 2)))
    return out


def batched_block(autorift, steps, npairs=256, npoints=1_000_000, seed=5):
    """The uneven batched workload: `npairs` pairs with sides uniform in [1024, 8192], each the
    reference warped by its own affine field (|d| <= 15 px, bilinear) plus N(0, 2^2) noise, made
    on the device; `npoints` scattered fractional centres spread evenly over the pairs, chip size
    from {16, 32, 48, 64} and search radius uniform in 8..40 per point; one batched pass
    (64x refinement) captured in a HIP graph and replayed."""
    import torch
    import torch.nn.functional as F
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    k1 = torch.arange(-6, 7, dtype=torch.float32, device=dev)
    k1 = torch.exp(-k1 * k1 / (2 * 1.5 ** 2))
    k1 = k1 / k1.sum()
    bt = autorift.Batch(autorift.params())
    sides = rng.integers(1024, 8193, (npairs, 2))
    fields = []
    t0 = time.time()
    for k in range(npairs):
        ny, nx = int(sides[k, 0]), int(sides[k, 1])
        # column-major ny x nx image = row-major tensor (nx, ny): dim 0 is x, dim 1 is y
        t = torch.randn((1, 1, nx, ny), generator=gen, device=dev)
        # separable Gaussian blur as sums of shifted slices (elementwise kernels; a convolution
        # library call would tune itself anew for every image shape)
        tp = F.pad(t, (0, 0, 6, 6), mode="reflect")
        t = sum(k1[j] * tp[:, :, j:j + nx, :] for j in range(13))
        tp = F.pad(t, (6, 6, 0, 0), mode="reflect")
        t = sum(k1[j] * tp[:, :, :, j:j + ny] for j in range(13))
        del tp
        t = (t - t.min()) * (255.0 / (t.max() - t.min()))
        # affine field (u, v) in pixels about the centre; |d| <= 15 over the image
        tr = rng.uniform(-9, 9, 2)
        lin = rng.uniform(-1, 1, (2, 2)) * (5.0 / max(nx, ny))
        cx, cy = (nx - 1) / 2, (ny - 1) / 2
        xs = torch.arange(nx, dtype=torch.float32, device=dev).view(nx, 1)
        ys = torch.arange(ny, dtype=torch.float32, device=dev).view(1, ny)
        u = tr[0] + lin[0, 0] * (xs - cx) + lin[0, 1] * (ys - cy)
        v = tr[1] + lin[1, 0] * (xs - cx) + lin[1, 1] * (ys - cy)
        # secondary(x, y) = reference(x + u, y + v): the tracker reports (dx, dy) = (u, v)
        gx = (ys + v) * (2.0 / (ny - 1)) - 1.0          # grid_sample x runs along dim 1 (y)
        gy = (xs + u) * (2.0 / (nx - 1)) - 1.0
        grid = torch.stack([gx.expand(nx, ny), gy.expand(nx, ny)], dim=-1).unsqueeze(0)
        sec = F.grid_sample(t, grid, mode="bilinear", padding_mode="border", align_corners=True)
        sec = sec + 2.0 * torch.randn(sec.shape, generator=gen, device=dev)
        ref = t.contiguous()
        sec = sec.contiguous()
        bt.add_pair_device(ref.data_ptr(), sec.data_ptr(), ny, nx)
        fields.append((tr, lin, cx, cy))
        del t, ref, sec, grid, u, v, gx, gy
        if (k + 1) % 32 == 0:
            log(f"batched workload: {k + 1} of {npairs} pairs made and prepared "
                f"({time.time() - t0:.1f}s)")
    torch.cuda.synchronize()
    gen_s = time.time() - t0
    # points spread evenly over the pairs, fractional 1-based centres
    pair = rng.integers(0, npairs, npoints).astype(np.int32)
    ny_of, nx_of = sides[pair, 0], sides[pair, 1]
    x = 1 + rng.random(npoints) * (nx_of - 1)
    y = 1 + rng.random(npoints) * (ny_of - 1)
    cs = rng.choice(np.array([16, 32, 48, 64]), npoints).astype(np.int64)
    r = rng.integers(8, 41, npoints).astype(np.int64)
    flds = [torch.from_numpy(np.ascontiguousarray(a)).to(dev)
            for a in (x, y, r, r, np.zeros(npoints), np.zeros(npoints), cs, cs)]
    pr = torch.from_numpy(pair).to(dev)
    out = [torch.empty(npoints, dtype=torch.float32, device=dev) for _ in range(3)]
    out.append(torch.empty(npoints, dtype=torch.uint8, device=dev))
    fp = [f.data_ptr() for f in flds]
    op = [o.data_ptr() for o in out]
    log(f"batched workload: {npoints} points; tuning the radius split")
    t1 = time.time()
    bt.track_device(fp, pr.data_ptr(), npoints, 64, 0, op)       # tunes the radius split
    tune_s = time.time() - t1
    split = bt.info()["split"]
    log(f"batched workload: split {split} chosen in {tune_s:.1f}s; capturing the pass")
    bt.track_device(fp, pr.data_ptr(), npoints, 64, 0, op, graph=True, split=split)  # capture
    torch.cuda.synchronize()
    ms = []
    t2 = time.perf_counter()
    for _ in range(steps):
        bt.track_device(fp, pr.data_ptr(), npoints, 64, 0, op, graph=True, split=split)
        ms.append(bt.info()["ms"])
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t2) / steps
    info = bt.info()
    dx = out[0].cpu().numpy()
    dy = out[1].cpu().numpy()
    searched = int(out[3].sum().item())
    ok = np.isfinite(dx)
    # expected (u, v) at each point (0-based pixel coordinates)
    tr = np.array([f[0] for f in fields])[pair]
    lin = np.array([f[1] for f in fields])[pair]
    cxx = np.array([f[2] for f in fields])[pair]
    cyy = np.array([f[3] for f in fields])[pair]
    ex = tr[:, 0] + lin[:, 0, 0] * (x - 1 - cxx) + lin[:, 0, 1] * (y - 1 - cyy)
    ey = tr[:, 1] + lin[:, 1, 0] * (x - 1 - cxx) + lin[:, 1, 1] * (y - 1 - cyy)
    err = np.hypot(dx[ok] - ex[ok], dy[ok] - ey[ok])
    flop = 2.0 * (2 * r) * (2 * r) * cs.astype(np.float64) ** 2
    # chips of area 2304 (48 x 48): the default pass correlates the unrounded mean-removed chip;
    # the exact pass (direct Float64 tracker, bit-equal to the oracle) is timed and compared
    def grab():
        torch.cuda.synchronize()
        return out[0].cpu().numpy(), out[1].cpu().numpy()
    bt.track_device(fp, pr.data_ptr(), npoints, 1, 0, op, split=split)
    d1 = grab()
    bt.set_exact(True)
    bt.track_device(fp, pr.data_ptr(), npoints, 1, 0, op, split=split)
    e1 = grab()
    bt.track_device(fp, pr.data_ptr(), npoints, 64, 0, op, split=split)
    exact_ms = bt.info()["ms"]
    e64 = grab()
    bt.close()
    m48 = (cs == 48) & np.isfinite(d1[0]) & np.isfinite(e1[0])
    peaks = float(np.mean((d1[0][m48] == e1[0][m48]) & (d1[1][m48] == e1[1][m48])))
    b48 = (cs == 48) & ok & np.isfinite(e64[0])
    off = np.hypot(dx[b48] - e64[0][b48], dy[b48] - e64[1][b48])
    chip48 = {"points": int(m48.sum()), "integer_peak_agreement": peaks,
              "subpixel_within_0_05px": float(np.mean(off <= 0.05)),
              "subpixel_offset_p999_px": float(np.quantile(off, 0.999)),
              "subpixel_offset_max_px": float(off.max()),
              "exact_pass_ms": exact_ms,
              "exact_points_per_s": searched / (exact_ms / 1000)}
    log(f"batched chip 48: integer peaks equal {peaks:.5f} over {m48.sum()} points, subpixel "
        f"within 0.05 px {chip48['subpixel_within_0_05px']:.5f}; exact pass {exact_ms:.1f} ms")
    ms_med = float(np.median(ms))
    res = {"pairs": npairs, "points": npoints, "searched": searched,
           "measured": int(ok.sum()), "ms_per_pass": ms_med, "wall_ms_per_pass": wall * 1000,
           "points_per_s": searched / (ms_med / 1000),
           "alg_tflop_per_s": float(flop.sum() / (ms_med / 1000) / 1e12),
           "radius_classes": list(split), "launches": info["launches"],
           "graph_replayed": info["replayed"],
           "median_error_px": float(np.median(err)), "p95_error_px": float(np.percentile(err, 95)),
           "generate_s": gen_s, "tune_s": tune_s, "chip48_vs_exact": chip48}
    log(f"batched: {npairs} pairs, {npoints} points, {ms_med:.1f} ms/pass "
        f"({res['points_per_s'] / 1e6:.1f} M points/s, {res['alg_tflop_per_s']:.0f} TFLOP/s "
        f"algorithmic), split {split}, launches {info['launches']}, median error "
        f"{res['median_error_px']:.3f} px (generate {gen_s:.1f}s, tune {tune_s:.1f}s)")
    return res


def profiled_traffic():
    """HBM bytes per tracker launch from the committed counter summary (None if absent)."""
    try:
        with open(os.path.join(ROOT, TRAFFIC_PROFILE)) as f:
            for line in f:
                if line.startswith("traffic:"):
                    return int(line.split()[1])
    except OSError:
        pass
    return None


def copy_calibration(nbytes):
    """A plain device copy of known size, for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE."""
    import torch
    x = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes}), flush=True)


REF_FINE_POINTS_PER_S = 5060.0   # 900 points / 177.86 ms (min), one Apple M2 Max thread, Julia
                                 # 1.12: "track/fine c32 r25 1024x1024" in the reference's
                                 # benchmark/results/baseline.json:35-38
REF_THREADED_POINTS_PER_S = 36320.0   # 900 points / 24.78 ms (min) on 8 M2 Max threads:
                                      # "track/threaded c32 r25 1024x1024", baseline.json:197-199


def host_cpus():
    """(threads this process may use, nproc, CPU model): the affinity mask, capped by the cgroup
    CPU quota when one is set."""
    import subprocess
    aff = len(os.sched_getaffinity(0))
    use = aff
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()
        if quota != "max":
            use = max(1, min(aff, int(float(quota) / float(period))))
    except (OSError, ValueError):
        pass
    model = None
    try:
        for ln in subprocess.run(["lscpu"], capture_output=True, text=True,
                                 timeout=10).stdout.splitlines():
            if ln.startswith("Model name:"):
                model = ln.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    return use, os.cpu_count(), model


def cpu_baseline(ref8, sec8, pts, n, npts, sample):
    """The oracle's tracker (FFT numerator, as the reference uses at chip 32 r 25) on the same
    UInt8 pair: serial on one thread and threaded on every usable core, each the minimum of 5
    timed runs after a warm-up, on evenly spaced samples of the grid sized for about 2 s a run."""
    import oracle
    threads, nproc, model = host_cpus()
    ok = np.ones((n, n), bool)
    names = ("x", "y", "radius_x", "radius_y", "chip_size_x", "chip_size_y")
    full = {f: np.ravel(getattr(pts, f), order="F") for f in names}

    def run(k, nth, reps=5):
        sel = np.linspace(0, npts - 1, k).astype(np.int64)
        f = {key: v[sel] for key, v in full.items()}
        t = []
        for _ in range(reps + 1):                  # the first run is the warm-up
            tc = time.perf_counter()
            oracle.track(ref8, sec8, ok, f["x"], f["y"], f["radius_x"], f["radius_y"],
                         f["chip_size_x"], f["chip_size_y"], up=64, num_mode=oracle.NUM_FFT,
                         nthreads=nth)
            t.append(time.perf_counter() - tc)
        return k / min(t[1:]), sum(t)

    k1 = min(sample or 1000, npts)
    serial, s1 = run(k1, 1)
    kt = min(sample or int(min(npts, max(1000, serial * threads * 2))), npts)
    threaded, s2 = run(kt, threads)
    cpu = {"value": threaded, "unit": "points/s", "cores": threads, "kind": "port",
           "sample": f"{kt} of the {npts} grid points (threaded) and {k1} (serial), evenly "
                     f"spaced; same UInt8 images; FFT numerator; min of 5 runs after a warm-up "
                     f"({s1 + s2:.1f} s in all)",
           "serial_points_per_s": serial, "threaded_points_per_s": threaded,
           "per_thread_points_per_s": threaded / threads, "nproc": nproc, "cpu_model": model,
           "reference_serial_points_per_s": REF_FINE_POINTS_PER_S,
           "reference_threaded_points_per_s_8_threads": REF_THREADED_POINTS_PER_S,
           "reference_source": "benchmark/results/baseline.json:35,197 (Apple M2 Max)",
           "reference_projected_points_per_s": REF_FINE_POINTS_PER_S * threads,
           "projection": "the reference's one-thread rate times the cores used here (ideal "
                         "scaling; an estimate of the reference on this host, not a run)"}
    log(f"cpu baseline: serial {serial:.0f} points/s, {threads} threads {threaded:.0f} points/s "
        f"({threaded / threads:.0f} per thread; the reference's fine pass: "
        f"{REF_FINE_POINTS_PER_S:.0f} points/s on one M2 thread); {model}, nproc {nproc}")
    return cpu


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=10000)
    ap.add_argument("--spacing", type=int, default=24)
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="points per CPU baseline run (0: ~1000 serial, ~2 s threaded)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--latency-pairs", type=int, default=20)
    ap.add_argument("--serve-rates", default="1,2,5,10,20,50,100,200",
                    help="Poisson arrival rates (pairs/s) of the serving sweep; empty: skip")
    ap.add_argument("--serve-requests", type=int, default=60,
                    help="most requests per rate (3 s of arrivals, at least 10)")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-pool", type=int, default=6)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-served-compare", action="store_true",
                    help="skip the served BF16 against plain Float32 comparison")
    ap.add_argument("--no-coherence", action="store_true")
    ap.add_argument("--no-cfg2", action="store_true",
                    help="skip the full multi-chip search on the production-size pair")
    ap.add_argument("--batch-pairs", type=int, default=256,
                    help="pairs of the batched workload (0: skip it)")
    ap.add_argument("--batch-points", type=int, default=1_000_000)
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only run a device copy of this many bytes (profiling calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 "
                                   "(sharding across GPUs is designed, not run)"}))
        sys.exit(2)

    import torch

    import autorift

    n = args.size
    t0 = time.time()
    tex, sec, fx, fy = synthetic_pair(n, fields=True)
    log(f"synthetic pair {n}x{n} in {time.time() - t0:.1f}s")
    grid = autorift.gridpoints((n, n), args.spacing, chip_size=128, search_radius=25)
    pts = grid.rebuild(chip_size_x=np.full(grid.shape, 32, np.int64),
                       chip_size_y=np.full(grid.shape, 32, np.int64))
    npts = len(pts)
    ctx = autorift.Context(autorift.params())
    t0 = time.perf_counter()
    ctx.set_images(tex, sec)          # default pipeline on the device: highpass(5) + UInt8
    prep_s = time.perf_counter() - t0
    ref8, sec8, _, _ = ctx.get_pair()
    log(f"device preparation (upload + highpass + quantize, both images) {prep_s:.2f}s")

    dev = torch.device("cuda:0")
    fields = [torch.from_numpy(np.ascontiguousarray(np.ravel(getattr(pts, f), order="F"))).to(dev)
              for f in ("x", "y", "radius_x", "radius_y", "dx_prior", "dy_prior", "chip_size_x",
                        "chip_size_y")]
    out = [torch.empty(npts, dtype=torch.float32, device=dev) for _ in range(3)]
    out.append(torch.empty(npts, dtype=torch.uint8, device=dev))
    fptr = [t.data_ptr() for t in fields]
    optr = [t.data_ptr() for t in out]
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        ctx.track_device(fptr, npts, 64, 0, optr)
    torch.cuda.synchronize()
    kms, bms, tms, rms = [], [], [], []
    t1 = time.perf_counter()
    for _ in range(args.steps):
        ctx.track_device(fptr, npts, 64, 0, optr)
        kms.append(ctx.last_kernel()[0])
        b, t, r = ctx.last_timing()
        bms.append(b)
        tms.append(t)
        rms.append(r)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    ms_step = (t2 - t1) * 1000 / args.steps
    searched = int(out[3].sum().item())
    value = searched / (ms_step / 1000)
    kavg, tavg, ravg = float(np.mean(kms)), float(np.mean(tms)), float(np.mean(rms))
    bavg = float(np.mean(bms))
    # the matrix-core roofline belongs to the correlation kernel (HIP events on its stream)
    achieved = searched * FLOP_PER_POINT / (tavg / 1000) / 1e12
    dx = out[0].cpu().numpy()
    log(f"{npts} points, searched {searched}, {ms_step:.2f} ms/step, kernels {kavg:.2f} ms "
        f"(box_u8_kernel {bavg:.2f}, track_u8_kernel {tavg:.2f}, refine_kernel {ravg:.2f}), "
        f"measured {np.count_nonzero(~np.isnan(dx))}, path {ctx.last_kernel()[1]}")

    mixed = None
    if not args.no_mixed:
        mixed = mixed_block(autorift, tex, sec, pts, min(args.steps, 3))
    cfg2 = None
    if not args.no_cfg2 and n == 10000:
        cfg2 = cfg2_block(autorift, tex, sec, fx, fy)
    del tex, sec, fx, fy

    coherence = None
    if not args.no_coherence:
        coherence = coherence_block(autorift, min(args.steps, 3))

    batched = None
    if args.batch_pairs > 0:
        batched = batched_block(autorift, min(args.steps, 5), args.batch_pairs, args.batch_points)

    served_vs_plain = None
    if not args.no_mixed and not args.no_served_compare:
        served_vs_plain = served_mixed_vs_plain(autorift)

    serving = None
    if args.serve_rates:
        rates = [float(r) for r in args.serve_rates.split(",") if r]
        serving = serving_sweep(autorift, rates, args.serve_requests, args.serve_slots,
                                args.serve_pool)

    lat = None
    if args.latency_pairs > 0:
        lat = pair_latency(autorift, args.latency_pairs)
        log(f"pair latency p50 {lat['p50_ms']:.2f} ms p99 {lat['p99_ms']:.2f} ms "
            f"({lat['pairs_per_s']:.1f} pairs/s), median dx {lat['median_dx']:.3f} "
            f"dy {lat['median_dy']:.3f}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ref8, sec8, pts, n, npts, args.cpu_sample)

    # the counter summary covers the default workload only
    traffic = profiled_traffic() if (n, args.spacing) == (10000, 24) else None
    line = {
        "metric": "NCC search points/sec (chip 32, r=25), %roofline; p50/p99 pair latency",
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": f"track pass, {n}x{n} UInt8 pair, {npts} grid points "
                               f"(spacing {args.spacing}), chip 32, r 25, PyramidRefine(64)",
                   "points": npts},
        "kernels_ms": {"box_u8_kernel": bavg, "track_u8_kernel": tavg, "refine_kernel": ravg},
        "roofline": {"kernel": "track_u8_kernel", "bound": "mfma", "achieved": achieved,
                     "peak": INT8_PEAK_TOPS,
                     "unit": "TFLOP/s", "frac": achieved / INT8_PEAK_TOPS, "traffic": traffic,
                     "traffic_source": TRAFFIC_PROFILE if traffic is not None else None},
        "cpu_baseline": cpu,
        "pair_latency": lat,
        "serving": serving,
        "batched": batched,
        "mixed_precision": mixed,
        "coherence": coherence,
        "served_mixed_vs_plain_fp32": served_vs_plain,
        "multichip_10000sq": cfg2,
        "prepare_s": prep_s,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
