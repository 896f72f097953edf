"""Intersection throughput benchmark (one JSON line on stdout).

Workload: 2^24 rays against the SAH BVH of a seeded noise-displaced icosphere
of 1,310,720 triangles (subdivision level 8). Two ray sets live in HBM before
timing starts: coherent (pinhole camera, 4096 x 4096 pixel-centre rays,
generated by the camera kernel) and incoherent (origins uniform in the scene
box, directions uniform on the sphere). One step = four launches over the
whole set: closest-hit and any-hit on each ray set, i.e. 4 * 2^24 BVH queries.

`value` is queries per second over the timed steps (Mrays/s). The roofline
object is computed for the slowest of the four kernels from its algorithmic
bytes per launch (32 B ray + 16 B hit per ray, 32 B per box test, 36 B per
triangle test, 64 B per sphere test, counted by the kernel's counting mode)
over its HIP-event-timed average duration. The CPU baseline is the oracle
(C++ restatement of the reference traversal) on a bounded random sample of
the same four query sets, using every host core available to this process.

Further objects ride on the same line: `time_to_relmse` (the door-gap interior
at 1920 x 1080, depth 10: the measured wall time of progressive renders until
relMSE against a reference first reaches 0.02, neural-guided against the plain
path with the reference's Russian roulette and at its best threshold);
`cornell` (the Cornell box configuration on the GPU and the oracle's CPU
render of it on every host core); `path_tracer` (the plain GPU path tracer at
1920 x 1080, 64 spp, depth 8: samples/s, Mrays/s, the traversal launches'
algorithmic bytes over their HIP-event time against the HBM roofline, and the
oracle's CPU samples/s); `guiding` (per-sample variance and seconds per sample
of the estimators, the time to relMSE 0.02 they predict, and the measured
renders behind `time_to_relmse`); `batch` (64 seeded scenes rendered by one
hipGraph-replayed wavefront loop against one call per scene); `mlp` and
`integrators`.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mrays/s + time-to-relMSE≤0.02 at 1080p, neural-guided vs plain GPU path"
HBM_PEAK_GBS = 8000.0
LEVEL = 8
N_RAYS = 1 << 24


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def alg_bytes(n_rays: int, st) -> float:
    return 32.0 * n_rays + 16.0 * n_rays + 32.0 * st.nodes_visited + 36.0 * st.tris_tested + 64.0 * st.spheres_tested


def path_alg_bytes(st) -> float:
    """Algorithmic bytes of a render's traversal launches: 32 B ray + 16 B hit per query, 32 B per box
    test, 36 B per triangle test, 64 B per sphere test (counted by the launches themselves)."""
    return alg_bytes(int(st.rays), st)


def traversal_roofline(ctx, render) -> dict:
    """The traversal roofline of a render: `render()` runs it once and returns its bg_stats; it runs
    twice more, once counting every traversal launch's box and primitive tests (the algorithmic
    bytes, path_alg_bytes) and once with a HIP event pair around each traversal launch."""
    ctx.set_profiling(count=True)
    cst = render()
    ctx.set_profiling(traversal_time=True)
    st = render()
    ctx.set_profiling()
    trav = path_alg_bytes(cst)
    return {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", "traversal_ms": round(st.traverse_ms, 2),
            "traversal_bytes": trav, "traversal_achieved": round(trav / st.traverse_ms / 1e6, 1),
            "traversal_frac": round(trav / st.traverse_ms / 1e6 / HBM_PEAK_GBS, 4),
            "box_tests": int(cst.nodes_visited), "triangle_tests": int(cst.tris_tested),
            "sphere_tests": int(cst.spheres_tested), "queries": int(cst.rays)}


TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "k_trav_pairs_intersection_1.31Mtri_2p24rays.md")


def traffic_from_profile(path: str, query: str):
    """HBM bytes per launch of one query's traversal kernel as the committed rocprofv3 summary of this
    workload gives them: its FETCH_SIZE and WRITE_SIZE table (KiB per dispatch) scaled by the copy
    calibration factors stated beside it. None if the summary is absent or lacks the row."""
    import re
    if not os.path.exists(path):
        return None
    text = open(path).read()
    ff = re.search(r"FETCH_SIZE \(factor ([\d.]+)\)", text)
    wf = re.search(r"WRITE_SIZE \(factor ([\d.]+)\)", text)
    label = query.replace("_", " ").replace(" hit", "")  # "incoherent_closest" -> "incoherent closest"
    row = re.search(r"^\| " + re.escape(label) + r" \| ([\d,]+) \| [\d.]+ \| ([\d,]+) ->", text, flags=re.M)
    if not (ff and wf and row):
        return None
    f, w = (float(row.group(k).replace(",", "")) for k in (1, 2))
    return round((f * float(ff.group(1)) + w * float(wf.group(1))) * 1024.0, 0)


def traffic_from(prof_dir: str, any_hit: bool, set_index: int):
    """HBM bytes per launch of the timed traversal kernel for one query, from rocprofv3 --pmc runs of
    `bench.py --no-cpu --steps 1 --warmup 0` (FETCH_SIZE and WRITE_SIZE, KiB per dispatch) scaled by the
    factors a plain copy of known size gives (`bench.py --copy-bytes N`)."""
    import csv
    import glob

    def rows(pattern):
        files = glob.glob(os.path.join(prof_dir, pattern))
        return list(csv.DictReader(open(files[0]))) if files else []

    import re
    # timed traversal instantiations: <BLOCK, STACK, ANY, COUNT=false, ...>
    tag = re.compile(r"k_trav\w*<\d+, \d+, " + ("true" if any_hit else "false") + ", false")

    def kernel_values(pattern, counter):
        vals = [float(r["Counter_Value"]) for r in rows(pattern)
                if r["Counter_Name"] == counter and tag.search(r["Kernel_Name"])]
        return vals[set_index] if len(vals) > set_index else None

    def copy_factor(pattern, counter):
        vals = [float(r["Counter_Value"]) for r in rows(pattern)
                if r["Counter_Name"] == counter and "copyBuffer" in r["Kernel_Name"]]
        if not vals:
            return None
        return (1 << 30) / 1024.0 / (sum(vals) / len(vals))

    f = kernel_values("prof_fetch/*counter_collection.csv", "FETCH_SIZE")
    w = kernel_values("prof_write/*counter_collection.csv", "WRITE_SIZE")
    ff = copy_factor("prof_copyf/*counter_collection.csv", "FETCH_SIZE")
    wf = copy_factor("prof_copyw/*counter_collection.csv", "WRITE_SIZE")
    if None in (f, w, ff, wf):
        return None
    return round((f * ff + w * wf) * 1024.0, 0)


def cpu_baseline(scene, bvh, sets, target_s: float):
    """Times the oracle's reference-order traversal on a random sample of each query set."""
    sys.path.insert(0, str(ROOT / "tests"))
    import cpu_oracle as O

    cores = len(os.sched_getaffinity(0))
    o = O.OracleScene(scene)
    nodes, slots = bvh.bvh()
    o.use_bvh(nodes, slots)
    rng = np.random.default_rng(1234)
    # calibrate on a small sample, then size the sample to about target_s of CPU time
    probe = 1 << 13
    t0 = time.perf_counter()
    for name, rays, any_hit in sets:
        o.intersect(rays[rng.integers(0, rays.shape[0], probe)], any_hit=any_hit, threads=cores)
    dt = time.perf_counter() - t0
    per_set = int(min(N_RAYS, max(probe, probe * target_s / max(dt, 1e-6))))
    samples = [(rays[rng.integers(0, rays.shape[0], per_set)], any_hit) for _, rays, any_hit in sets]
    t0 = time.perf_counter()
    outs = [o.intersect(r, any_hit=a, threads=cores) for r, a in samples]
    dt = time.perf_counter() - t0
    return {
        "value": round(per_set * len(sets) / dt / 1e6, 3),
        "unit": "Mrays/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{per_set} random rays from each of the 4 query sets ({per_set * len(sets)} queries, {dt:.1f} s)",
    }, samples, outs


def path_tracer_leg(ctx, spp: int, cpu: bool) -> dict:
    """The plain GPU path tracer at 1920 x 1080: 2,048,000 triangles + 256 spheres, 4 quad lights,
    `spp` samples per pixel (64 in the configuration), max depth 8, corrected (pbrt-v3) semantics."""
    from bonsai_amd import GpuBvh
    from bonsai_amd.integrator import MODE_CORRECTED, MODE_REF, Film, GpuPathIntegrator
    from bonsai_amd.scenes import path_tracer_scene

    scene = path_tracer_scene()
    bvh = GpuBvh(ctx, scene)
    out = {"workload": "plain path tracer 1920x1080, max depth 8", "triangles": scene.n_tris,
           "spheres": scene.n_spheres, "spp": spp}
    for mode, name in ((MODE_CORRECTED, "corrected"), (MODE_REF, "reference_semantics")):
        gi = GpuPathIntegrator(ctx, scene.camera, spp=spp, max_depth=8, mode=mode, seed=5)
        film = Film(ctx, scene.camera.width, scene.camera.height)
        gi.render_into(bvh, film, spp=1)  # warm-up (allocations, code objects)
        # untimed: the box / primitive tests of every traversal launch (the algorithmic bytes)
        ctx.set_profiling(count=True)
        film.clear()
        cst = gi.render_into(bvh, film)
        # timed, with a HIP event pair around each traversal kernel
        ctx.set_profiling(traversal_time=True)
        film.clear()
        torch_sync()
        t0 = time.perf_counter()
        st = gi.render_into(bvh, film)
        wall = time.perf_counter() - t0
        ctx.set_profiling()
        trav = path_alg_bytes(cst)
        whole = trav + 16.0 * st.samples
        out[name] = {"seconds": round(wall, 3), "samples_per_s": round(st.samples / wall, 1),
                     "Msamples_s": round(st.samples / wall / 1e6, 2), "rays": int(st.rays),
                     "Mrays_s": round(st.rays / wall / 1e6, 1), "shadow_rays": int(st.shadow_rays),
                     "kernel_ms": round(st.kernel_ms, 1), "image_mean": float(f"{film.image().mean():.6g}"),
                     "roofline": {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "traversal_ms": round(st.traverse_ms, 2),
                                  "traversal_bytes": trav,
                                  "traversal_achieved": round(trav / st.traverse_ms / 1e6, 1),
                                  "traversal_frac": round(trav / st.traverse_ms / 1e6 / HBM_PEAK_GBS, 4),
                                  "render_bytes": whole,
                                  "render_achieved": round(whole / st.kernel_ms / 1e6, 1),
                                  "render_frac": round(whole / st.kernel_ms / 1e6 / HBM_PEAK_GBS, 4),
                                  "box_tests": int(cst.nodes_visited), "triangle_tests": int(cst.tris_tested),
                                  "sphere_tests": int(cst.spheres_tested), "queries": int(cst.rays)}}
        assert cst.rays == st.rays
        log(f"path tracer ({name}): {spp} spp in {wall:.2f} s, {st.rays / wall / 1e6:.0f} Mrays/s, traversal "
            f"{st.traverse_ms:.1f} ms = {trav / st.traverse_ms / 1e6:.0f} GB/s algorithmic")
        film.close()
    if cpu:
        c = cpu_path_rate(scene, bvh, 1, 8, MODE_CORRECTED)
        c["gpu_vs_cpu"] = round(out["corrected"]["samples_per_s"] / c["samples_per_s"], 1)
        out["cpu_baseline"] = c
        log(f"path tracer, CPU oracle: {c['Msamples_s']} Msamples/s on {c['cores']} cores")
    bvh.close()
    return out


def cpu_path_rate(scene, bvh, spp: int, depth: int, mode: int, rr: float = 1.0, target_s: float = 6.0) -> dict:
    """The oracle's path integrator (the reference's rayon path restated in C++, every host core) on a
    horizontal band of the film through its middle, sized to about target_s of CPU time."""
    sys.path.insert(0, str(ROOT / "tests"))
    import cpu_oracle as O

    cores = len(os.sched_getaffinity(0))
    o = O.OracleScene(scene)
    nodes, slots = bvh.bvh()
    o.use_bvh(nodes, slots)
    o.set_shading(scene)
    W, H = scene.camera.width, scene.camera.height

    def band(rows, n_spp):
        y0 = max(0, H // 2 - rows // 2)
        t0 = time.perf_counter()
        o.render(scene, n_spp, depth, mode=mode, seed=5, rr=rr, window=(0, y0, W, y0 + rows), threads=cores)
        return time.perf_counter() - t0

    # grow the band (then, with the whole film, the samples per pixel) until a render takes at least
    # half the target: a small probe's time is mostly fixed cost, so one extrapolation falls short
    rows = 16
    dt = band(rows, spp)
    for _ in range(6):
        if dt >= 0.5 * target_s:
            break
        f = min(64.0, target_s / max(dt, 1e-3))
        if rows < H:
            rows = int(min(H, max(rows + 1, rows * f)))
        else:
            spp = max(spp + 1, int(spp * f))
        dt = band(rows, spp)
    n = W * rows * spp
    return {"samples_per_s": round(n / dt, 1), "Msamples_s": round(n / dt / 1e6, 4), "cores": cores,
            "kind": "port", "sample": f"{W}x{rows} band through the film's middle at {spp} spp ({dt:.1f} s)"}


def cornell_leg(ctx, cpu: bool) -> dict:
    """The plain CPU path tracer's configuration: Cornell box 256 x 256, 16 spp, max depth 5, the quad
    light one-sided, in both semantics: the GPU path (median of 5 renders) and, beside it, the oracle's
    full render of the same configuration on every host core."""
    from bonsai_amd import GpuBvh
    from bonsai_amd.integrator import MODE_CORRECTED, MODE_REF, Film, GpuPathIntegrator
    from bonsai_amd.scenes import cornell_box

    s = cornell_box()
    bvh = GpuBvh(ctx, s)
    out = {"workload": "Cornell box 256x256, 16 spp, max depth 5: 10 wall + 2 light triangles, 2 spheres"}
    for mode, name in ((MODE_REF, "reference_semantics"), (MODE_CORRECTED, "corrected")):
        gi = GpuPathIntegrator(ctx, s.camera, spp=16, max_depth=5, mode=mode, seed=1)
        film = Film(ctx, 256, 256)
        gi.render_into(bvh, film)
        times = []
        for _ in range(5):
            film.clear()
            torch_sync()
            t0 = time.perf_counter()
            st = gi.render_into(bvh, film)
            times.append(time.perf_counter() - t0)
        film.close()
        wall = float(np.median(times))
        entry = {"gpu_seconds": round(wall, 5), "gpu_Msamples_s": round(st.samples / wall / 1e6, 1),
                 "gpu_Mrays_s": round(st.rays / wall / 1e6, 1), "rays_per_sample": round(st.rays / st.samples, 3)}
        if cpu:
            sys.path.insert(0, str(ROOT / "tests"))
            import cpu_oracle as O
            cores = len(os.sched_getaffinity(0))
            o = O.OracleScene(s)
            nodes, slots = bvh.bvh()
            o.use_bvh(nodes, slots)
            o.set_shading(s)
            t0 = time.perf_counter()
            o.render(s, 16, 5, mode=mode, seed=1, threads=cores)
            dt = time.perf_counter() - t0
            entry["cpu_baseline"] = {"seconds": round(dt, 3), "Msamples_s": round(st.samples / dt / 1e6, 3),
                                     "cores": cores, "kind": "port", "sample": "the full configuration"}
            entry["gpu_vs_cpu"] = round(dt / wall, 1)
        out[name] = entry
        log(f"Cornell ({name}): GPU {wall * 1e3:.2f} ms" + (f", CPU oracle {entry['cpu_baseline']['seconds']:.2f} s"
                                                            if cpu else ""))
    bvh.close()
    return out


def integrators_leg(ctx, spp: int) -> dict:
    """The reference's other integrators and the two-lobe material on the path-tracer scene at
    1920 x 1080: DirectLightingIntegrator with UniformSampleAll and UniformSampleOne, Whitted, and
    the path tracer with every non-black matte material turned plastic (ks 0.25, roughness 0.2),
    all with corrected semantics; then the film's GPU resolve (Film::write_image) and the OpenEXR
    write of the result."""
    import tempfile

    from bonsai_amd import GpuBvh
    from bonsai_amd import _native as N
    from bonsai_amd.integrator import (MODE_CORRECTED, UNIFORM_SAMPLE_ALL, UNIFORM_SAMPLE_ONE,
                                       DirectLightingIntegrator, Film, GpuPathIntegrator, WhittedIntegrator)
    from bonsai_amd.scenes import path_tracer_scene, plastic

    scene = path_tracer_scene()
    cam = scene.camera
    out = {"workload": "path-tracer scene 1920x1080 (2,048,000 triangles + 256 spheres, 8 light triangles)",
           "spp": spp}
    bvh = GpuBvh(ctx, scene)
    runs = [("direct_all_lights", DirectLightingIntegrator(ctx, cam, spp, strategy=UNIFORM_SAMPLE_ALL,
                                                           mode=MODE_CORRECTED, seed=5)),
            ("direct_one_light", DirectLightingIntegrator(ctx, cam, spp, strategy=UNIFORM_SAMPLE_ONE,
                                                          mode=MODE_CORRECTED, seed=5)),
            ("whitted", WhittedIntegrator(ctx, cam, spp, mode=MODE_CORRECTED, seed=5))]
    film = Film(ctx, cam.width, cam.height)
    for name, integ in runs:
        integ.render_into(bvh, film, spp=1)
        film.clear()
        t0 = time.perf_counter()
        st = integ.render_into(bvh, film)
        wall = time.perf_counter() - t0
        out[name] = {"seconds": round(wall, 3), "Msamples_s": round(st.samples / wall / 1e6, 2),
                     "rays": int(st.rays), "Mrays_s": round(st.rays / wall / 1e6, 1),
                     "shadow_rays": int(st.shadow_rays)}
        log(f"{name}: {spp} spp in {wall:.2f} s, {st.rays / wall / 1e6:.0f} Mrays/s")
    # Film::write_image on the GPU and the OpenEXR file
    d_ms = []
    import torch

    buf = torch.empty(cam.width * cam.height * 3, dtype=torch.float32, device="cuda:0")
    for _ in range(5):
        ms = C.c_double(0.0)
        N.check(N.lib().bg_film_resolve_device(film.h, 1.0, 1.0, C.c_void_p(buf.data_ptr()), C.byref(ms)), ctx.h)
        d_ms.append(ms.value)
    with tempfile.TemporaryDirectory() as td:
        t0 = time.perf_counter()
        film.write_image(os.path.join(td, "result.exr"))
        exr_s = time.perf_counter() - t0
        size = os.path.getsize(os.path.join(td, "result.exr"))
    npx = cam.width * cam.height
    res_ms = float(np.median(d_ms[1:]))
    out["film"] = {"resolve_kernel_ms": round(res_ms, 4), "resolve_GB_s": round(28 * npx / res_ms / 1e6, 1),
                   "write_image_exr_s": round(exr_s, 4), "exr_bytes": size}
    log(f"film resolve {res_ms * 1e3:.1f} us, write_image + EXR {exr_s * 1e3:.1f} ms ({size} bytes)")
    film.close()
    bvh.close()
    # the two-lobe material: non-black mattes turned plastic
    mats = []
    for m in scene.materials:
        if m.type == N.BG_MAT_MATTE and any(m.reflectance[k] > 0 for k in range(3)):
            m = plastic(tuple(m.reflectance), (0.25, 0.25, 0.25), 0.2)
        mats.append(m)
    scene.materials = mats
    bvh = GpuBvh(ctx, scene)
    gi = GpuPathIntegrator(ctx, cam, spp=spp, max_depth=8, mode=MODE_CORRECTED, seed=5)
    film = Film(ctx, cam.width, cam.height)
    gi.render_into(bvh, film, spp=1)
    film.clear()
    t0 = time.perf_counter()
    st = gi.render_into(bvh, film)
    wall = time.perf_counter() - t0
    out["path_plastic"] = {"seconds": round(wall, 3), "Msamples_s": round(st.samples / wall / 1e6, 2),
                           "rays": int(st.rays), "Mrays_s": round(st.rays / wall / 1e6, 1)}
    log(f"path tracer, plastic: {spp} spp in {wall:.2f} s, {st.rays / wall / 1e6:.0f} Mrays/s")
    film.close()
    bvh.close()
    return out


def box_blur(img: np.ndarray, r: int) -> np.ndarray:
    """Mean over the (2r+1)^2 neighbourhood, clipped at the image border."""
    out = img.astype(np.float64)
    for axis in (0, 1):
        c = np.cumsum(np.pad(out, [(r + 1, r) if a == axis else (0, 0) for a in range(out.ndim)], mode="edge"),
                      axis=axis)
        n = out.shape[axis]
        hi = np.take(c, np.arange(2 * r + 1, 2 * r + 1 + n), axis=axis)
        lo = np.take(c, np.arange(0, n), axis=axis)
        out = (hi - lo) / (2 * r + 1)
    return out


def batch_leg(ctx, spp: int) -> dict:
    """64 seeded scenes of uneven size (triangle budgets log-uniform in 1k-500k, square films of side
    uniform in 128-1024), 16 spp, depth 6, mixed materials: one batched wavefront loop (hipGraph
    replay of a captured pass, autotuned traversal launch shape) against one render call per scene."""
    from bonsai_amd import GpuBvh
    from bonsai_amd.integrator import MODE_CORRECTED, SCHEDULE_PERSISTENT, BatchRenderer, Film, GpuPathIntegrator
    from bonsai_amd.scenes import batch_scene

    scenes = [batch_scene(i) for i in range(64)]
    bvhs = [GpuBvh(ctx, s) for s in scenes]
    items = [(b, s.camera, 1000 + k) for k, (b, s) in enumerate(zip(bvhs, scenes))]
    pixels = sum(s.camera.width * s.camera.height for s in scenes)
    out = {"workload": "64 scenes, 16 spp, depth 6", "scenes": 64, "triangles": sum(s.n_tris for s in scenes),
           "pixels": pixels, "spp": spp}
    br = BatchRenderer(ctx, items)
    t0 = time.perf_counter()
    br.render(1, 6, mode=MODE_CORRECTED, autotune=True)  # tunes the launch shape; warms allocations
    tune_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    st = br.render(spp, 6, mode=MODE_CORRECTED, autotune=False)
    wall = time.perf_counter() - t0
    b, s_, c, v = br.launch()
    out["batched"] = {"seconds": round(wall, 3), "Msamples_s": round(st.samples / wall / 1e6, 2),
                      "Mrays_s": round(st.rays / wall / 1e6, 1), "launch": {"block": b, "stack": s_,
                      "blocks_per_cu": c, "variant": v}, "first_call_seconds_incl_autotune": round(tune_s, 2)}
    # (measured renders run their passes without the graph)
    out["batched"]["roofline"] = traversal_roofline(ctx, lambda: br.render(spp, 6, mode=MODE_CORRECTED,
                                                                            autotune=False))
    # the same batch with each pass as one persistent path kernel (path slots handed to free lanes
    # from an atomic counter; the same films): scheduling measured against the wavefront passes
    br.render(1, 6, mode=MODE_CORRECTED, autotune=False, schedule=SCHEDULE_PERSISTENT)
    torch_sync()
    t0 = time.perf_counter()
    stp = br.render(spp, 6, mode=MODE_CORRECTED, autotune=False, schedule=SCHEDULE_PERSISTENT)
    wallp = time.perf_counter() - t0
    out["persistent"] = {"seconds": round(wallp, 3), "Msamples_s": round(stp.samples / wallp / 1e6, 2),
                         "Mrays_s": round(stp.rays / wallp / 1e6, 1), "wavefront_speedup": round(wallp / wall, 3)}
    log(f"batch leg: persistent path kernel {wallp:.2f} s against {wall:.2f} s for the wavefront passes")
    br.close()
    films = [Film(ctx, s.camera.width, s.camera.height) for s in scenes]
    for (bvh, cam, seed), f in zip(items, films):  # warm-up
        GpuPathIntegrator(ctx, cam, spp=1, max_depth=6, mode=MODE_CORRECTED, seed=seed).render_into(bvh, f)
    torch_sync()
    t0 = time.perf_counter()
    rays = 0
    for (bvh, cam, seed), f in zip(items, films):
        f.clear()
        rays += GpuPathIntegrator(ctx, cam, spp=spp, max_depth=6, mode=MODE_CORRECTED, seed=seed).render_into(bvh, f).rays
    wall1 = time.perf_counter() - t0
    out["per_scene"] = {"seconds": round(wall1, 3), "Msamples_s": round(pixels * spp / wall1 / 1e6, 2),
                        "Mrays_s": round(rays / wall1 / 1e6, 1)}
    out["speedup_batched_vs_per_scene"] = round(wall1 / wall, 3)
    rf = out["batched"]["roofline"]
    log(f"batch leg: batched {wall:.2f} s, per-scene {wall1:.2f} s; batched traversal {rf['traversal_ms']:.1f} ms = "
        f"{rf['traversal_achieved']:.0f} GB/s algorithmic")
    for f in films:
        f.close()
    for b_ in bvhs:
        b_.close()
    return out


MFMA_BF16_PEAK_TFLOPS = 2500.0
MLP_MACS = 4 * 64 * 64 + 64 * 40  # multiply-adds per sample and forward pass


def mlp_leg(ctx) -> dict:
    """Guiding network on device-resident batches: inference over 2^22 samples and Adam steps on
    2^16-sample batches. Forward = 2 * 18,944 flop per sample; a training step = forward + input
    and weight gradients = 3 forward's worth. Against the dense BF16 matrix-core peak."""
    import torch
    from bonsai_amd.guide import GuideMLP, init_params

    n_inf, n_tr = 1 << 22, 1 << 16
    g = GuideMLP(ctx, init_params(3), n_tr)
    gen = torch.Generator(device="cpu").manual_seed(5)
    x = torch.rand((n_inf, 12), generator=gen).to("cuda:0")
    out = torch.empty((n_inf, 40), dtype=torch.float32, device="cuda:0")
    tg = torch.nn.functional.normalize(torch.randn((n_tr, 4), generator=gen), dim=1).to("cuda:0")
    tg[:, 3] = 1.0
    torch_sync()
    g.eval_device(x.data_ptr(), n_inf, out.data_ptr())
    inf_ms = min(g.eval_device(x.data_ptr(), n_inf, out.data_ptr()) for _ in range(5))
    g.train_device(x.data_ptr(), tg.data_ptr(), n_tr)
    tr_ms = min(g.train_device(x.data_ptr(), tg.data_ptr(), n_tr) for _ in range(5))
    inf_tf = 2 * MLP_MACS * n_inf / inf_ms / 1e9
    tr_tf = 3 * 2 * MLP_MACS * n_tr / tr_ms / 1e9
    g.close()
    return {"inference": {"samples": n_inf, "ms": round(inf_ms, 3), "Msamples_s": round(n_inf / inf_ms / 1e3, 1),
                          "TFLOP_s": round(inf_tf, 1), "mfma_frac": round(inf_tf / MFMA_BF16_PEAK_TFLOPS, 4)},
            "train_step": {"batch": n_tr, "ms": round(tr_ms, 3), "TFLOP_s": round(tr_tf, 1),
                           "mfma_frac": round(tr_tf / MFMA_BF16_PEAK_TFLOPS, 4)},
            "peak_TFLOP_s": MFMA_BF16_PEAK_TFLOPS}


def rel_mse_pair(a: np.ndarray, b: np.ndarray) -> float:
    """relMSE estimated from two independent renders of n samples each: per pixel (A - B)^2 / 2 is an
    unbiased estimate of one render's variance, and the squared mean in the denominator is taken
    from the 7 x 7 box-filtered average of both renders (a single noisy pixel in the denominator would
    hide its own error). For unbiased estimators with per-sample relative variance v this is v / n."""
    m = box_blur(0.5 * (a + b), 3)
    return float(np.mean(0.5 * (a - b) ** 2 / (m * m + 0.01)))


# Settings of the guided estimators on the door-gap workload, from sweeps there (predicted time to
# relMSE 0.02, warm-up included): Russian roulette off (with the reference's rule, q = 1 - max(beta)
# after three bounces, roulette removes the deep low-throughput paths that reach the light through
# the gap and the survivors' 1/(1-q) boosts dominate the error: guided 91 s with it, 45 s without);
# the network guides the first 4 vertices (3: 29 s, 4: 26 s, 5: 26 s, 6: 50 s, all: 52 s); it is
# trained during a warm-up render of 3x the measured samples with 48 Adam steps per 2-spp pass
# (warm-up spp and steps: 32 and 8: 45 s; 64 and 16: 32 s; 64 and 32: 23 s; 128 and 16: 26 s;
# 128 and 32: 19.4 s; 64 and 64: 18.6 s; 96 and 48: 18.2 s; 48 and 96: 19.8 s; 256 and 32: 26 s).
GUIDE_BOUNCES = 4
GUIDE_RR = 0.0
GUIDE_STEPS = 48
GUIDE_WARMUP_FACTOR = 3
# The plain path's best roulette threshold on the same workload, from {1 (the reference's default),
# 0.25, 0.1, 0.05, 0.02, 0}: 114, 107, 98, 89, 91 and 92 s predicted.
PLAIN_BEST_RR = 0.05


RELMSE_TARGET = 0.02


def rel_mse_terms(img: np.ndarray, ref: np.ndarray) -> float:
    """relMSE = mean over pixels and RGB of (I - R)^2 / (R^2 + 0.01), FP64."""
    return float(np.mean((img - ref) ** 2 / (ref * ref + 0.01)))


def measured_time_to_relmse(ctx, scene, bvh, ref_spp: int, max_seconds: float) -> dict:
    """Progressive 1920 x 1080 renders of the door-gap interior, each checked against a reference
    image after every step (the check is not timed): the wall time at which relMSE = mean (I - R)^2 /
    (R^2 + 0.01) first reaches RELMSE_TARGET. The reference R is the frozen-network guided estimator
    (unbiased, with the lowest variance per second) at ref_spp samples per pixel in two independent
    halves R1, R2 with FP64 film sums; its own error, mean ((R1 - R2) / 2)^2 / (R^2 + 0.01), is
    subtracted from every measurement (I and R are independent, so E (I - R)^2 = Var I + Var R).
    Plain estimators step 256 spp at a time, the guided one 64 spp after its timed warm-up."""
    from bonsai_amd.guide import GuideMLP, init_params
    from bonsai_amd.integrator import MODE_CORRECTED, Film, GpuPathIntegrator, GuidedPathIntegrator, resolve

    W, H = scene.camera.width, scene.camera.height

    def guided(g, seed, train, n):
        return GuidedPathIntegrator(ctx, scene.camera, spp=n, max_depth=10, guide=g, seed=seed, select_prob=0.5,
                                    steps=GUIDE_STEPS, train=train, bounces=GUIDE_BOUNCES, rr_threshold=GUIDE_RR)

    def warm_up(g, seed):
        film = Film(ctx, W, H)
        torch_sync()
        t0 = time.perf_counter()
        n = GUIDE_WARMUP_FACTOR * 32
        for k in range(0, n, 2):
            guided(g, seed, True, 2).render_into(bvh, film, spp=2, first_sample=(1 << 20) + k)
        torch_sync()
        film.close()
        return time.perf_counter() - t0

    def image(film):
        rgb, w = film.sums(f64=True)
        return resolve(rgb, w)

    # the reference
    gref = GuideMLP(ctx, init_params(11), 65536)
    warm_up(gref, 800)
    halves = []
    for seed in (801, 802):
        film = Film(ctx, W, H, f64=True)
        integ = guided(gref, seed, False, ref_spp // 2)
        for k in range(0, ref_spp // 2, 256):
            integ.render_into(bvh, film, spp=min(256, ref_spp // 2 - k), first_sample=k)
        halves.append(image(film))
        film.close()
    R = 0.5 * (halves[0] + halves[1])
    ref_err = float(np.mean(0.25 * (halves[0] - halves[1]) ** 2 / (R * R + 0.01)))
    out = {"reference": {"estimator": "guided, frozen network", "spp": ref_spp, "relmse_of_reference": round(ref_err, 5)},
           "target": RELMSE_TARGET, "check_excluded_from_time": True}
    log(f"measured relMSE: reference {ref_spp} spp, its own relMSE {ref_err:.4f}")

    def progressive(name, make, step, warm_s):
        film = Film(ctx, W, H, f64=True)
        t, n, hist = warm_s, 0, []
        reached = None
        while t < max_seconds:
            integ = make(step)
            torch_sync()
            t0 = time.perf_counter()
            integ.render_into(bvh, film, spp=step, first_sample=n)
            torch_sync()
            t += time.perf_counter() - t0
            n += step
            e = rel_mse_terms(image(film), R) - ref_err
            hist.append((n, round(t, 3), round(e, 5)))
            log(f"  {name}: {n} spp, {t:.1f} s, relMSE {e:.4f}")
            if e <= RELMSE_TARGET:
                reached = (n, t, e)
                break
        film.close()
        rec = {"step_spp": step, "warmup_seconds": round(warm_s, 3), "checks": hist[-4:]}
        if reached:
            n1, t1, e1 = reached
            # between checks relMSE falls as 1/spp: the sample count and time where it crosses the target
            n0, t0_, e0 = hist[-2] if len(hist) > 1 else (0, warm_s, float("inf"))
            n_star = n1 * e1 / RELMSE_TARGET
            t_star = t0_ + (t1 - t0_) * (n_star - n0) / (n1 - n0) if n1 > n0 else t1
            rec.update({"seconds_at_first_check_below": round(t1, 2), "spp_at_first_check_below": n1,
                        "seconds": round(min(t1, max(t_star, t0_)), 2), "spp": int(round(n_star))})
        else:
            n1, t1, e1 = hist[-1]
            rec.update({"seconds": None, "not_reached_within_s": max_seconds, "relmse_at_end": e1,
                        "seconds_extrapolated": round(warm_s + (t1 - warm_s) * e1 / RELMSE_TARGET, 1)})
        log(f"measured relMSE {name}: " + (f"{rec['seconds']} s ({rec['spp']} spp)" if reached else
                                          f"not reached in {max_seconds} s, relMSE {e1:.4f}"))
        return rec

    def plain(rr, seed):
        return lambda n: GpuPathIntegrator(ctx, scene.camera, spp=n, max_depth=10, mode=MODE_CORRECTED, seed=seed,
                                           rr_threshold=rr)

    out["plain"] = progressive("plain", plain(1.0, 811), 256, 0.0)
    out["plain"]["rr_threshold"] = 1.0
    out["plain_best_rr"] = progressive("plain_best_rr", plain(PLAIN_BEST_RR, 812), 256, 0.0)
    out["plain_best_rr"]["rr_threshold"] = PLAIN_BEST_RR
    g = GuideMLP(ctx, init_params(3), 65536)
    warm_s = warm_up(g, 813)
    out["guided"] = progressive("guided", lambda n: guided(g, 814, False, n), 64, warm_s)
    gd = out["guided"]["seconds"]
    for key, ref in (("speedup_guided_vs_plain", "plain"), ("speedup_guided_vs_plain_best_rr", "plain_best_rr")):
        p = out[ref]["seconds"]
        out[key] = round(p / gd, 3) if (p and gd) else None
    gref.close()
    g.close()
    return out


def guiding_leg(ctx, spp: int, measure: bool = True, ref_spp: int = 2048, max_seconds: float = 150.0,
                cpu: bool = True) -> dict:
    """Door-gap interior at 1920 x 1080, depth 10: per-sample relative variance v (from two
    independent renders of `spp` samples each) and seconds per sample of each estimator; the
    predicted time to relMSE 0.02 is v * seconds_per_spp / 0.02 (Monte Carlo error falls as
    1 / samples), plus any training warm-up. "plain" is the plain path with the reference's default
    Russian roulette, "plain_best_rr" with its best roulette threshold on this workload;
    "no_network" is the guided estimator's structure with the network left out; "guided_training"
    trains its network from scratch during both measured renders; "guided" trains it during a
    warm-up render and measures with the network frozen. The guided estimators use the settings
    above (GUIDE_*)."""
    from bonsai_amd import GpuBvh
    from bonsai_amd.guide import GuideMLP, init_params
    from bonsai_amd.integrator import MODE_CORRECTED, Film, GpuPathIntegrator, GuidedPathIntegrator
    from bonsai_amd.scenes import door_gap_room

    scene = door_gap_room()
    bvh = GpuBvh(ctx, scene)
    W, H = scene.camera.width, scene.camera.height
    out = {"workload": "door-gap interior 1920x1080, max depth 10, 1 light behind a 5 cm gap",
           "triangles": scene.n_tris, "spp_per_render": spp, "guided_bounces": GUIDE_BOUNCES,
           "guided_rr_threshold": GUIDE_RR, "guided_adam_steps_per_pass": GUIDE_STEPS,
           "warmup_spp": GUIDE_WARMUP_FACTOR * spp, "plain_best_rr_threshold": PLAIN_BEST_RR}
    guide = GuideMLP(ctx, init_params(3), 65536)

    def render(g, sel, train, seed, rr, n=spp, first=0):
        # training renders call once per 2 samples (the network trains after every call); the
        # others render all samples in one call, in the library's passes (about 64M paths)
        if g is None:
            gi = GpuPathIntegrator(ctx, scene.camera, spp=n, max_depth=10, mode=MODE_CORRECTED, seed=seed,
                                   rr_threshold=rr)
        else:
            gi = GuidedPathIntegrator(ctx, scene.camera, spp=n, max_depth=10, guide=g, seed=seed, select_prob=sel,
                                      steps=GUIDE_STEPS, train=train, bounces=GUIDE_BOUNCES, rr_threshold=rr)
        film = Film(ctx, W, H)
        rays = 0
        torch_sync()
        t0 = time.perf_counter()
        step = 2 if train else n
        for k in range(0, n, step):
            st = gi.render_into(bvh, film, spp=min(step, n - k), first_sample=first + k)
            rays += st.rays
        secs = time.perf_counter() - t0
        img = film.image()
        film.close()
        return img, secs, rays

    variants = [("plain", None, 0.0, False, False, 1.0), ("plain_best_rr", None, 0.0, False, False, PLAIN_BEST_RR),
                ("no_network", guide, 0.0, False, False, GUIDE_RR),
                ("guided_training", guide, 0.5, True, False, GUIDE_RR), ("guided", guide, 0.5, False, True, GUIDE_RR)]
    for name, g, sel, train, warm, rr in variants:
        if g is not None:
            g.set_params(init_params(3))
        warm_s = 0.0
        if warm:
            _, warm_s, _ = render(g, sel, True, 30, rr, n=GUIDE_WARMUP_FACTOR * spp, first=1 << 20)
        res = [render(g, sel, train, seed, rr) for seed in (31, 32)]
        secs = sum(r[1] for r in res)
        rays = sum(r[2] for r in res)
        v = rel_mse_pair(res[0][0], res[1][0]) * spp
        t_spp = secs / (2 * spp)
        out[name] = {"rr_threshold": rr, "v_relmse_x_spp": round(v, 3), "seconds_per_spp": round(t_spp, 4),
                     "Mrays_s": round(rays / secs / 1e6, 1),
                     "predicted_seconds_to_relmse_0.02": round(warm_s + v * t_spp / 0.02, 1)}
        if warm:
            out[name]["warmup_training_seconds"] = round(warm_s, 3)
        if name in ("plain", "guided"):  # (a training render would change the network it measures)
            def one(g=g, sel=sel, rr=rr):
                if g is None:
                    gi = GpuPathIntegrator(ctx, scene.camera, spp=8, max_depth=10, mode=MODE_CORRECTED, seed=33,
                                           rr_threshold=rr)
                else:
                    gi = GuidedPathIntegrator(ctx, scene.camera, spp=8, max_depth=10, guide=g, seed=33,
                                              select_prob=sel, steps=GUIDE_STEPS, train=False,
                                              bounces=GUIDE_BOUNCES, rr_threshold=rr)
                film = Film(ctx, W, H)
                st = gi.render_into(bvh, film)
                film.close()
                return st
            out[name]["roofline_8spp"] = traversal_roofline(ctx, one)
        log(f"guiding leg {name}: v {v:.2f}, {t_spp * 1e3:.1f} ms/spp, {rays / secs / 1e6:.0f} Mrays/s"
            + (f", warm-up {warm_s:.2f} s" if warm else ""))
    gd = out["guided"]["predicted_seconds_to_relmse_0.02"]
    for key, ref in (("speedup_guided_vs_plain", "plain"), ("speedup_guided_vs_plain_best_rr", "plain_best_rr")):
        p = out[ref]["predicted_seconds_to_relmse_0.02"]
        out[key] = round(p / gd, 3) if gd > 0 else None
    if cpu:
        c = cpu_path_rate(scene, bvh, 1, 10, MODE_CORRECTED)
        c["gpu_vs_cpu_plain"] = round(1.0 / out["plain"]["seconds_per_spp"] * W * H / c["samples_per_s"], 1)
        out["cpu_baseline"] = c
        log(f"door gap, CPU oracle: {c['Msamples_s']} Msamples/s on {c['cores']} cores")
    if measure:
        out["measured"] = measured_time_to_relmse(ctx, scene, bvh, ref_spp, max_seconds)
    bvh.close()
    return out


def torch_sync() -> None:
    import torch
    torch.cuda.synchronize()


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--render-spp", type=int, default=64, help="path-tracer leg samples per pixel (0: skip)")
    ap.add_argument("--guide-spp", type=int, default=32, help="guiding leg samples per render (0: skip)")
    ap.add_argument("--batch-spp", type=int, default=16, help="scene-batch leg samples per pixel (0: skip)")
    ap.add_argument("--no-mlp", action="store_true", help="skip the guiding-network leg")
    ap.add_argument("--no-relmse", action="store_true",
                    help="skip the measured time-to-relMSE renders of the guiding leg (the prediction stays)")
    ap.add_argument("--relmse-ref-spp", type=int, default=2048, help="reference samples per pixel (two halves)")
    ap.add_argument("--relmse-max-seconds", type=float, default=150.0,
                    help="longest progressive render per estimator before giving up")
    ap.add_argument("--no-cornell", action="store_true", help="skip the Cornell-box (CPU path tracer) leg")
    ap.add_argument("--integrators-spp", type=int, default=16,
                    help="direct / Whitted / plastic leg samples per pixel (0: skip)")
    ap.add_argument("--skip-intersection", action="store_true",
                    help="run only the other legs (profiling them alone); value is then null")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--stack", type=int, default=0)
    ap.add_argument("--blocks-per-cu", type=int, default=-1)
    ap.add_argument("--variant", type=int, default=-1, help="0: per-ray rounds, 1: per-lane refill")
    ap.add_argument("--traffic-from", default="",
                    help="directory holding rocprofv3 --pmc outputs prof_fetch/, prof_write/, prof_copyf/, prof_copyw/ "
                         "of this workload; fills roofline.traffic")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only copy a buffer of this many bytes device-to-device three times (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against a known traffic)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        return 2

    import torch

    if args.copy_bytes:
        src = torch.ones(args.copy_bytes // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": args.copy_bytes, "copies": 3}), flush=True)
        return 0

    from bonsai_amd import Context, GpuBvh
    from bonsai_amd.scenes import incoherent_rays, intersection_camera, intersection_scene

    ctx = Context(0)
    if args.block or args.stack or args.blocks_per_cu >= 0 or args.variant >= 0:
        b, s, c, v = ctx.launch()
        ctx.set_launch(args.block or b, args.stack or s, c if args.blocks_per_cu < 0 else args.blocks_per_cu,
                       v if args.variant < 0 else args.variant)
    value, wall, breakdown, roofline, cpu, scene = None, 0.0, None, None, None, None
    queries = []
    if not args.skip_intersection:
        t0 = time.perf_counter()
        scene = intersection_scene(LEVEL)
        gen_s = time.perf_counter() - t0
        bvh = GpuBvh(ctx, scene)
        log(f"scene: {scene.n_tris} tris, {bvh.info.n_nodes} nodes, height {bvh.info.max_depth}, "
            f"build+upload {bvh.info.build_ms:.0f} ms, generation {gen_s * 1e3:.0f} ms")

        dev = torch.device("cuda:0")
        cam = intersection_camera(4096, 4096)
        coh = torch.empty((N_RAYS, 8), dtype=torch.float32, device=dev)
        ctx.camera_rays_device(cam, coh.data_ptr())
        lo, hi = scene.bounds()
        inc_host = incoherent_rays(N_RAYS, lo, hi, 99)
        inc = torch.from_numpy(inc_host).to(dev)
        hits = torch.empty((N_RAYS, 4), dtype=torch.int32, device=dev)
        torch.cuda.synchronize()
        queries = [("coherent_closest", coh, False), ("coherent_any", coh, True),
                   ("incoherent_closest", inc, False), ("incoherent_any", inc, True)]

        # algorithmic bytes per launch from the kernel's counting mode (untimed)
        counts = {}
        for name, rays, any_hit in queries:
            st = bvh.intersect_device(rays.data_ptr(), N_RAYS, hits.data_ptr(), any_hit, count=True)
            counts[name] = (alg_bytes(N_RAYS, st), st.nodes_visited / N_RAYS, st.tris_tested / N_RAYS)
            log(f"{name}: {st.nodes_visited / N_RAYS:.1f} box tests/ray, {st.tris_tested / N_RAYS:.1f} tri tests/ray")

        def step(record):
            for name, rays, any_hit in queries:
                st = bvh.intersect_device(rays.data_ptr(), N_RAYS, hits.data_ptr(), any_hit)
                if record is not None:
                    record.setdefault(name, []).append(st.kernel_ms)

        for _ in range(args.warmup):
            step(None)
        torch.cuda.synchronize()
        kms: dict = {}
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step(kms)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        total = N_RAYS * len(queries) * args.steps
        value = total / wall / 1e6
        breakdown = {}
        for name, _, _ in queries:
            ms = float(np.mean(kms[name]))
            b = counts[name][0]
            breakdown[name] = {"Mrays_s": round(N_RAYS / ms / 1e3, 1), "kernel_ms": round(ms, 3),
                               "alg_GB_s": round(b / ms / 1e6, 1), "box_tests_per_ray": round(counts[name][1], 2),
                               "tri_tests_per_ray": round(counts[name][2], 2)}
        dom = max(breakdown, key=lambda k: breakdown[k]["kernel_ms"])
        achieved = breakdown[dom]["alg_GB_s"]
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": (traffic_from(args.traffic_from, dom.endswith("_any"), 1 if dom.startswith("incoherent") else 0)
                                if args.traffic_from else traffic_from_profile(TRAFFIC_PROFILE, dom)),
                    "traffic_source": (f"rocprofv3 --pmc runs in {args.traffic_from}" if args.traffic_from else
                                       "profiles/" + os.path.basename(TRAFFIC_PROFILE)),
                    "kernel": f"k_traverse ({dom})",
                    "bytes_per_launch": counts[dom][0]}

        cpu = None
        if not args.no_cpu:
            sets = [(name, (coh if name.startswith("coherent") else None), a) for name, _, a in queries]
            coh_host = coh.cpu().numpy()
            sets = [(n, coh_host if n.startswith("coherent") else inc_host, a) for n, _, a in sets]
            cpu, samples, outs = cpu_baseline(scene, bvh, sets, args.cpu_seconds)
            # the sampled GPU results must equal the oracle's
            for (r, a), ref in zip(samples, outs):
                g = (bvh.intersect_test(r) if a else bvh.intersect(r)).view(np.uint32).reshape(-1, 4)
                if not np.array_equal(g[:, 0], ref[:, 0]):
                    log("PARITY FAILURE on the CPU-baseline sample")
                    cpu["parity"] = False
                    break
            else:
                cpu["parity"] = True
    b, s_, c, v_ = ctx.launch()
    hits = coh = inc = None
    torch.cuda.empty_cache()
    cl = cornell_leg(ctx, not args.no_cpu) if not args.no_cornell else None
    pt = path_tracer_leg(ctx, args.render_spp, not args.no_cpu) if args.render_spp > 0 else None
    gl = (guiding_leg(ctx, args.guide_spp, not args.no_relmse, args.relmse_ref_spp, args.relmse_max_seconds,
                      not args.no_cpu) if args.guide_spp > 0 else None)
    ttr = None
    if gl and "measured" in gl:
        m = gl["measured"]
        ttr = {"workload": "door-gap interior 1920x1080, depth 10", "target_relmse": RELMSE_TARGET,
               "guided_s": m["guided"]["seconds"], "plain_s": m["plain"]["seconds"],
               "plain_best_rr_s": m["plain_best_rr"]["seconds"],
               "speedup_vs_plain": m["speedup_guided_vs_plain"],
               "speedup_vs_plain_best_rr": m["speedup_guided_vs_plain_best_rr"], "measured": True}
    bl = batch_leg(ctx, args.batch_spp) if args.batch_spp > 0 else None
    ml = mlp_leg(ctx) if not args.no_mlp else None
    il = integrators_leg(ctx, args.integrators_spp) if args.integrators_spp > 0 else None
    out = {
        "metric": METRIC,
        "value": round(value, 2) if value is not None else None,
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3) if value is not None else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "intersection microbenchmark: 2^24 coherent + 2^24 incoherent rays, closest-hit "
                               "and any-hit, SAH BVH over a seeded displaced icosphere",
                   "triangles": scene.n_tris if scene else None, "rays_per_set": N_RAYS,
                   "queries_per_step": N_RAYS * len(queries),
                   "launch": {"block": b, "stack": s_, "blocks_per_cu": c, "variant": v_}},
        "breakdown": breakdown,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "time_to_relmse": ttr,
        "cornell": cl,
        "path_tracer": pt,
        "guiding": gl,
        "batch": bl,
        "mlp": ml,
        "integrators": il,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
