"""Full-scan benchmark: FP32 forward + volume VJP + 6-DoF pose VJP.

Workload: 256^3 FP32 seeded ellipsoid phantom at 1 mm, circular orbit of 360
views (angles 2*pi*k/360), SOD/SDD 1000/1500 mm, 512x512 detector with the
full-support coverage pitch, 1024 clipped midpoint samples per ray.  One step
is one iteration as the reference benchmark defines it
(benchmarks/fullscan.py:224-229): a fresh Projector forward at pose tangent
xi = 0, the squared-error residual against a seeded target image, and
torch.autograd.grad for the field and the FP64 pose together.

Prints one JSON line.  ``value`` = nominal samples per pass (360*512^2*1024)
divided by the median iteration time, in Gsamples/s.  ``roofline`` is for the
dominant kernel, timed with HIP events on its launch stream inside the timed
region: its measured HBM traffic (rocprofv3 FETCH_SIZE / WRITE_SIZE,
calibrated on a copy kernel, recorded for a digest of the kernel sources)
over its launch time against 8 TB/s, and its counted VALU instructions over
its launch time against the issue rate of independent FMA chains measured in
the same run.  ``cpu_baseline`` times the CPU oracle (a restatement of the
reference FP32 algorithm) on a bounded sample of the same workload.

Further blocks of the same line: the reference's component table, its
full-scan acceptance guards, its single-request case (chest AP view),
serving (batched BF16 / FP32 cells against the plain per-request form,
closed and open loop), the precision form, the exact cell-Gauss rule, the
deterministic gather (full scan and the reference's chest workload) and the
Monte Carlo transport with its oracle tally check.  ``--*-only`` flags run
one block alone; ``--kernels-once``, ``--copy-calibration`` and
``--record-counters`` drive and record the counter runs.
"""

from __future__ import annotations

import argparse
import csv
import gc
import hashlib
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
SHAPE = (256, 256, 256)
VIEWS = 360
DETECTOR = (512, 512)
SAMPLES = 1024
SOD, SDD = 1000.0, 1500.0
# algorithmic bytes per effective sample (FP32 storage): 8 corners x 4 B
GATHER_B = 32
SCATTER_B = 32


def _args():
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=5)
    parser.add_argument("--warmup", type=int, default=2)
    parser.add_argument("--views", type=int, default=VIEWS)
    parser.add_argument("--samples", type=int, default=SAMPLES)
    parser.add_argument("--storage", default="float32", choices=["float32", "bfloat16", "float16"])
    parser.add_argument("--no-cpu-baseline", action="store_true")
    parser.add_argument("--cpu-rows-stride", type=int, default=1)
    parser.add_argument("--no-serving", action="store_true", help="skip the batched pose-fit serving measurement")
    parser.add_argument("--serving-requests", type=int, default=200)
    parser.add_argument("--serving-rates", default="10,50",
                        help="Poisson arrival rates (requests/s) of the open-loop runs, comma separated")
    parser.add_argument("--serving-forms", default="batched_bf16,batched_fp32,plain_fp32",
                        help="serving forms for --serving-only (batched_bf16, batched_fp32, plain_fp32)")
    parser.add_argument("--serving-only", action="store_true",
                        help="only the serving measurement (every form at every rate; one JSON object)")
    parser.add_argument("--no-guards", action="store_true", help="skip the full-scan acceptance guards")
    parser.add_argument("--no-request", action="store_true", help="skip the single-request (chest AP) timings")
    parser.add_argument("--request-only", action="store_true", help="only the single-request (chest AP) timings")
    parser.add_argument("--request-repeats", type=int, default=300, help="timed repeats per single-request figure")
    parser.add_argument("--record-counters", nargs="+", default=None, metavar="CSV",
                        help="instead of the benchmark, write profiles/projector_counters.json from rocprofv3 "
                             "counter CSVs of --kernels-once runs (copy-calibration CSVs first, see "
                             "--calibration-bytes)")
    parser.add_argument("--calibration-bytes", type=float, default=0.0,
                        help="bytes read and written by the copy kernel of the calibration CSVs")
    parser.add_argument("--no-transport", action="store_true", help="skip the Monte Carlo transport measurement")
    parser.add_argument("--no-precision", action="store_true", help="skip the BF16/FP16 storage error report")
    parser.add_argument("--no-gather", action="store_true", help="skip the deterministic adjoint gather measurement")
    parser.add_argument("--no-cell-gauss", action="store_true", help="skip the exact cell-Gauss measurement")
    parser.add_argument("--cell-gauss-only", action="store_true",
                        help="only the exact cell-Gauss measurement (profiling driver)")
    parser.add_argument("--gather-chest-only", action="store_true",
                        help="only the deterministic gather on the chest workload (one JSON object)")
    parser.add_argument("--gather-only", action="store_true",
                        help="only the deterministic adjoint gather measurement (profiling driver)")
    parser.add_argument("--transport-only", action="store_true",
                        help="only the Monte Carlo transport measurement (profiling driver)")
    parser.add_argument("--transport-histories", type=int, default=100_000_000)
    parser.add_argument("--transport-cpu-histories", type=int, default=100_000_000)
    parser.add_argument(
        "--kernels-once",
        action="store_true",
        help="instead of the benchmark, launch forward, volume, pose and joint VJP once each (profiling driver)",
    )
    parser.add_argument(
        "--copy-calibration",
        type=float,
        default=0.0,
        help="instead of the benchmark, run a plain device copy of this many GiB (counter calibration)",
    )
    return parser.parse_args()


COUNTERS = ROOT / "profiles" / "projector_counters.json"


def sources_digest() -> str:
    """SHA-256 over the HIP sources and the C-ABI header: the code a counter
    profile measured."""
    h = hashlib.sha256()
    paths = sorted(p for p in (ROOT / "dpt" / "csrc").iterdir() if p.suffix in (".hip", ".h"))
    for path in [*paths, ROOT / "include" / "dpt_amd.h"]:
        h.update(path.name.encode())
        h.update(path.read_bytes())
    return h.hexdigest()


def _counter_rows(path):
    """{(dispatch, kernel): {counter: value}} of one rocprofv3 counter CSV."""
    out = {}
    with open(path) as f:
        for row in csv.DictReader(f):
            key = (int(row["Dispatch_Id"]), row["Kernel_Name"])
            out.setdefault(key, {})[row["Counter_Name"]] = float(row["Counter_Value"])
    return out


def record_counters(paths, calibration_bytes: float) -> int:
    """Write profiles/projector_counters.json from rocprofv3 counter CSVs.

    CSVs whose kernels include no projector kernel are the copy calibration:
    the largest kernel there moves ``calibration_bytes`` in and out, which
    scales FETCH_SIZE and WRITE_SIZE (KiB per dispatch) to bytes.  Projector
    kernels (forward32 / vjp32) keep their last dispatch per name; counters of
    several passes over the same driver are merged by kernel name."""
    fetch_scale = write_scale = 1024.0
    kernels = {}
    for path in paths:
        rows = _counter_rows(path)
        projector = {k: v for k, v in rows.items() if "forward32" in k[1] or "vjp32" in k[1]}
        if not projector:
            copy = max(rows.values(), key=lambda c: c.get("FETCH_SIZE", 0.0) + c.get("WRITE_SIZE", 0.0))
            if calibration_bytes > 0 and copy.get("FETCH_SIZE"):
                fetch_scale = calibration_bytes / copy["FETCH_SIZE"]
            if calibration_bytes > 0 and copy.get("WRITE_SIZE"):
                write_scale = calibration_bytes / copy["WRITE_SIZE"]
            continue
        for (_, name), values in sorted(projector.items()):
            kernels.setdefault(name, {}).update(values)
    for values in kernels.values():
        if "FETCH_SIZE" in values:
            values["fetch_bytes"] = values["FETCH_SIZE"] * fetch_scale
        if "WRITE_SIZE" in values:
            values["write_bytes"] = values["WRITE_SIZE"] * write_scale
    record = {
        "sources_sha256": sources_digest(),
        "fetch_bytes_per_kib_counted": fetch_scale,
        "write_bytes_per_kib_counted": write_scale,
        "kernels": kernels,
    }
    COUNTERS.write_text(json.dumps(record, indent=1, sort_keys=True) + "\n")
    print(json.dumps({"written": str(COUNTERS), "kernels": sorted(kernels)}), flush=True)
    return 0


def copy_calibration(gib: float) -> int:
    """A plain copy kernel of known size: HBM counters of this run calibrate
    FETCH_SIZE / WRITE_SIZE (bytes read = bytes written = the buffer size)."""
    import torch

    count = int(gib * (1 << 30)) // 4
    src = torch.ones(count, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": count * 4, "launches": 3}), flush=True)
    return 0


def kernels_once(args) -> int:
    """One launch each of the forward and the volume-only, pose-only and joint
    VJP kernels on the benchmark workload, through the C ABI (profiling driver)."""
    import numpy as np
    import torch

    from dpt import _native, phantoms

    grid = phantoms.centred_grid(SHAPE)
    values = phantoms.chest_phantom(grid, seed=2026)
    geometry, _ = phantoms.circular_geometry(grid, args.views, DETECTOR, sod=SOD, sdd=SDD)
    panels = np.asarray([p.packed() for p in geometry.views])
    field = torch.tensor(values, device="cuda:0")
    poses = torch.tensor([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], dtype=torch.float64, device="cuda:0")
    poses = poses.expand(args.views, 12).contiguous()
    out = torch.empty((args.views, *DETECTOR), device="cuda:0")
    grad_field = torch.empty(grid.shape, device="cuda:0")
    grad_pose = torch.empty((args.views, 12), dtype=torch.float64, device="cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    # the projector's default FP32 plan gathers from the cell layout
    # (DPT_AMD_GATHER=linear selects the linear gathers)
    cells = os.environ.get("DPT_AMD_GATHER") != "linear"
    needs = _native.VOLUME | _native.POSE | (_native.GATHER_CELLS if cells else 0)
    plan = _native.Plan(grid, views=args.views, rows=DETECTOR[0], cols=DETECTOR[1], panels=panels, rays=None,
                        samples=args.samples, arithmetic="float32", storage="float32", needs=needs, device=0)
    plan.forward(field.data_ptr(), poses.data_ptr(), out.data_ptr(), stream)
    seed = (out - 0.5 * out.mean()).contiguous()
    for flags in (_native.VOLUME, _native.POSE, _native.VOLUME | _native.POSE):
        plan.vjp(field.data_ptr(), poses.data_ptr(), seed.data_ptr(), grad_field.data_ptr(), grad_pose.data_ptr(),
                 flags, stream)
    torch.cuda.synchronize()
    codes = plan.status(stream)
    print(json.dumps({"kernels_once": True, "status_max": int(codes.max())}), flush=True)
    return 0


def serving_workload(count: int, seed: int = 404):
    """Pose-fit workload: 16 resident seeded phantoms with shapes drawn from
    {128^3 at 1 mm, 256^3 at 1 mm, 512x507x76 at 0.45x0.45x2.5 mm}; each
    request has 2-8 cone views at random angles, a 128^2/256^2/512^2 detector,
    256/512/768/1024 samples per ray, 20-120 iterations and a target rendered
    by the FP32 projector at a random pose (+/-10 deg, +/-20 mm) plus noise."""
    import math

    import numpy as np
    import torch

    from dpt import phantoms, serving
    from dpt.geometry import ConeBeamIntrinsics, RigidTransform, compose_pose
    from dpt.torch import Projector

    rng = np.random.default_rng(seed)
    shapes = [((128, 128, 128), (1.0, 1.0, 1.0)), ((256, 256, 256), (1.0, 1.0, 1.0)), ((76, 507, 512), (0.45, 0.45, 2.5))]
    volumes = []
    for i in range(16):
        shape, spacing = shapes[int(rng.integers(3))]
        grid = phantoms.centred_grid(shape, spacing)
        volumes.append((f"vol{i:02d}", grid, phantoms.chest_phantom(grid, seed=100 + i)))
    requests, truths = [], []
    on_device = {}
    for j in range(count):
        key, grid, values = volumes[int(rng.integers(16))]
        if key not in on_device:
            on_device[key] = torch.tensor(values, device="cuda:0")
        views = int(rng.integers(2, 9))
        det = int(rng.choice([128, 256, 512]))
        samples = int(rng.choice([256, 512, 768, 1024]))
        iterations = int(rng.integers(20, 121))
        angles = rng.uniform(0.0, 2 * math.pi, views)
        xi = np.concatenate([rng.uniform(-20, 20, 3), np.radians(rng.uniform(-10, 10, 3))])
        nx, ny, nz = grid.xyz
        diag = math.hypot(nx * grid.spacing_mm[0], ny * grid.spacing_mm[1], nz * grid.spacing_mm[2])
        sod = max(1000.0, 2.0 * diag)
        sdd = 1.5 * sod
        pitch = phantoms.coverage_pitch(grid, sod, sdd, angles, (det, det), margin=1.15)
        geometry = ConeBeamIntrinsics(sod, sdd, (det, det), (pitch, pitch), centre_mm=grid.centre_mm()).batch(angles)
        truth = compose_pose(RigidTransform(), tuple(xi)).packed()
        with torch.no_grad():
            image = Projector(grid=grid, geometry=geometry, samples_per_ray=samples)(
                on_device[key], packed_pose=torch.tensor(truth, dtype=torch.float64, device="cuda:0")
            )
            image += 1e-3 * torch.randn(image.shape, generator=torch.Generator("cuda:0").manual_seed(j), device="cuda:0")
        requests.append(serving.FitRequest(f"req{j:04d}", key, geometry, image.cpu().numpy(), samples, iterations))
        truths.append(truth)
    return volumes, requests, truths


SERVING_FORMS = {
    # name: (storage, layout, slots, hipGraph steps)
    "batched_bf16": ("bfloat16", "cells", 64, True),
    # ablation: the batched engine with the FP32 field in the cell layout
    "batched_fp32": ("float32", "cells", 64, True),
    # the reference's form: FP32 linear field, one solve at a time, direct launches
    "plain_fp32": ("float32", "linear", 1, False),
}


def serving_benchmark(count: int, rates, forms=("batched_bf16", "batched_fp32", "plain_fp32"),
                      open_forms=("batched_bf16",)) -> dict:
    """Pose-fit service: each form closed-loop (all requests queued at t = 0)
    and, for ``open_forms``, open-loop at each Poisson arrival rate in
    ``rates``; ``count`` requests per form and per rate.  Latency is arrival
    to completion; p50 / p99 over all requests of a run.  A progress line
    goes to stderr every 100 completions."""
    import dataclasses

    import numpy as np

    from dpt import serving

    volumes, requests, truths = serving_workload(count)
    out = {"requests_per_run": count, "arrival_rates_per_s": list(rates), "forms": {}}
    for name in forms:
        storage, layout, slots, graph = SERVING_FORMS[name]
        registry = serving.VolumeRegistry()
        for key, grid, values in volumes:
            registry.add(key, values, grid, storage=storage, layout=layout)
        engine = serving.PoseFitEngine(registry, max_requests=slots, max_tiles=1 << 19, use_graph=graph)
        engine.run(requests[:2], realtime=False)  # warm-up: code objects, graphs

        def progress(done, label=name):
            if done // 100 > progress.shown:
                progress.shown = done // 100
                _log(f"serving {label}: {done} done")

        progress.shown = 0
        closed = engine.run(requests, realtime=False, progress=progress)
        errors = np.asarray([serving.pose_error(r.pose, truths[int(r.key[3:])]) for r in closed.results])
        row = {"storage": storage, "layout": layout, "slots": slots, "graph": graph,
               "closed_loop": closed.summary(),
               "median_rotation_error_deg": round(float(np.median(errors[:, 0])), 5),
               "p99_rotation_error_deg": round(float(np.percentile(errors[:, 0], 99)), 5),
               "median_translation_error_mm": round(float(np.median(errors[:, 1])), 5)}
        _log(f"serving {name}: {row['closed_loop']['solves_per_s']} solves/s closed-loop")
        if name in open_forms:
            row["open_loop"] = {}
            for k, rate in enumerate(rates):
                at = serving.poisson_arrivals(count, rate, seed=405 + k)
                progress.shown = 0
                run = engine.run([dataclasses.replace(q, arrival=float(t)) for q, t in zip(requests, at)],
                                 realtime=True, progress=progress)
                row["open_loop"][f"{rate:g}_per_s"] = run.summary()
                _log(f"serving {name} at {rate:g} req/s: p50/p99 {run.summary()['p50_latency_ms']}/"
                     f"{run.summary()['p99_latency_ms']} ms")
        engine.close()
        out["forms"][name] = row
    if "batched_bf16" in out["forms"] and "plain_fp32" in out["forms"]:
        out["throughput_ratio_closed_loop"] = round(
            out["forms"]["batched_bf16"]["closed_loop"]["solves_per_s"]
            / out["forms"]["plain_fp32"]["closed_loop"]["solves_per_s"], 3)
    return out


# The reference's request-sized case (docs/performance.md, "Completed
# operation timings": chest AP, scalar, central cotangent, 512^2, 1024
# samples): FP32 volume VJP 0.958784 ms and pose VJP 0.580912 ms (CUDA
# events), forward + volume/pose VJP 3.834406 / 4.161676 ms median / p95
# completed wall; FP64 3.624160, 8.693632 and 20.512820 / 21.108925 ms.
REQUEST_REFERENCE_MS = {
    "fp32": {"volume_vjp": 0.958784, "pose_vjp": 0.580912, "iteration_p50": 3.834406, "iteration_p95": 4.161676},
    "fp64": {"volume_vjp": 3.624160, "pose_vjp": 8.693632, "iteration_p50": 20.512820, "iteration_p95": 21.108925},
}


def request_benchmark(repeats: int = 300) -> dict:
    """One request of the reference's shape: the chest grid (512 x 512 x 139
    at 0.762 x 0.762 x 2.5 mm; synthetic chest phantom content), its AP view on
    a 512^2 panel, S = 1024, the central-window cotangent (unit sum).  Volume
    and pose VJPs alone: torch events around autograd.grad on a retained
    graph; the iteration: forward + volume and pose VJP, completed wall time;
    medians and p95 over ``repeats``."""
    import numpy as np
    import torch

    from dpt import phantoms
    from dpt.geometry import DetectorGeometry
    from dpt.torch import Projector
    from dpt.volumes import GridSpec

    device = torch.device("cuda:0")
    grid = GridSpec(**CHEST_GRID)
    geometry = DetectorGeometry(**CHEST_AP)
    values = phantoms.chest_phantom(grid, seed=2027)
    rows, cols = CHEST_AP["shape"]
    result = {"workload": "chest AP view, 512 x 512 x 139 grid (synthetic content), 512^2 panel, S = 1024, "
                          "central-window cotangent", "repeats": repeats}
    for precision in ("fp32", "fp64"):
        module = Projector(grid=grid, geometry=geometry, precision=precision, samples_per_ray=1024)
        dtype = torch.float32 if precision == "fp32" else torch.float64
        seed = torch.zeros((1, rows, cols), dtype=dtype, device=device)
        seed[:, rows * 5 // 16: rows * 11 // 16, cols * 5 // 16: cols * 11 // 16] = 1.0
        seed /= seed.sum()
        field = torch.tensor(values, device=device, requires_grad=True)
        pose = torch.zeros(6, dtype=torch.float64, device=device, requires_grad=True)
        timings = {"volume_vjp": [], "pose_vjp": [], "iteration": []}
        for _ in range(3):  # warm-up
            torch.autograd.grad(module(field, pose), (field, pose), seed)
        torch.cuda.synchronize()
        for name, inputs in (("volume_vjp", (field,)), ("pose_vjp", (pose,))):
            image = module(field if name == "volume_vjp" else field.detach(),
                           pose if name == "pose_vjp" else pose.detach())
            for _ in range(repeats):
                start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                start.record()
                torch.autograd.grad(image, inputs, seed, retain_graph=True)
                stop.record()
                stop.synchronize()
                timings[name].append(start.elapsed_time(stop))
        for _ in range(repeats):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            torch.autograd.grad(module(field, pose), (field, pose), seed)
            torch.cuda.synchronize()
            timings["iteration"].append((time.perf_counter() - t0) * 1e3)
        ref = REQUEST_REFERENCE_MS[precision]
        row = {
            "volume_vjp_ms": round(float(np.median(timings["volume_vjp"])), 4),
            "pose_vjp_ms": round(float(np.median(timings["pose_vjp"])), 4),
            "iteration_p50_ms": round(float(np.median(timings["iteration"])), 4),
            "iteration_p95_ms": round(float(np.percentile(timings["iteration"], 95)), 4),
            "reference_ms": ref,
        }
        row["iteration_speedup_vs_reference"] = round(ref["iteration_p50"] / row["iteration_p50_ms"], 3)
        result[precision] = row
        _log(f"request {precision}: volume VJP {row['volume_vjp_ms']} ms, pose VJP {row['pose_vjp_ms']} ms, "
             f"iteration p50/p95 {row['iteration_p50_ms']}/{row['iteration_p95_ms']} ms "
             f"(reference {ref['iteration_p50']}/{ref['iteration_p95']})")
        del module, field, pose
    return result


def transport_benchmark(histories: int, cpu_histories: int, cpu: bool) -> dict:
    """Monte Carlo transport: 128^3 three-material chest label map, AP cone-beam
    point source, 256^2 primary/scatter tallies, 60 keV and a 40-120 keV
    spectrum.  Histories/s of simulate() (source sampling, Woodcock walk,
    tallies, moments, one status readback) and of the trace kernel alone."""
    import torch

    from dpt import transport

    out = {"workload": "128^3 chest labels (air/soft tissue/bone at 0.0012/1.0/1.85 g/cm^3), 1 mm, "
                       "SOD/SDD 1000/1500 mm, 256^2 detector, Klein-Nishina + Rayleigh, Woodcock majorant",
           "histories": histories}
    for name, spectrum in (("mono_60kev", False), ("spectrum_40_120kev", True)):
        spec, physical, labels, a, s, dens, source = transport.chest_transport_scene(spectrum=spectrum)
        ws = transport.prepare_transport(spec, material_ids=labels, absorption=a, scattering=s,
                                         max_histories=histories, physical=physical)
        sim = transport.prepare_simulation(ws)
        transport.simulate(dens, source=source, batch=transport.HistoryBatch(1, 0, min(1 << 20, histories)), workspace=sim)
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        start.record()
        result = transport.simulate(dens, source=source, batch=transport.HistoryBatch(2026, 0, histories), workspace=sim)
        stop.record()
        wall = time.perf_counter() - t0
        stop.synchronize()
        gpu_ms = start.elapsed_time(stop)
        primary = float(result.primary_mean.sum())
        scatter = float(result.scatter_mean.sum())
        out[name] = {
            "histories_per_s": round(histories / wall, 1),
            "wall_ms": round(wall * 1e3, 3),
            "gpu_ms": round(gpu_ms, 3),
            "status_histogram": list(result.status_histogram.counts),
            "detected_primary_per_history": round(primary, 6),
            "detected_scatter_per_history": round(scatter, 6),
            "scatter_to_primary": round(scatter / primary, 5) if primary > 0 else None,
        }
        if cpu:
            from oracle import transport as T

            from dpt.transport.simulation import bind_model

            model = T.Model(*bind_model(spec, material_ids=labels, absorption=a, scattering=s, physical=physical))
            t0 = time.perf_counter()
            cpu = T.trace(model, dens, source, 2026, 0, cpu_histories, records=False)
            elapsed = time.perf_counter() - t0
            # the same histories on the GPU: tallies and statuses against the oracle's
            if cpu_histories <= histories:
                import numpy as np

                same = transport.simulate(dens, source=source, batch=transport.HistoryBatch(2026, 0, cpu_histories),
                                          workspace=sim)
                mean, _ = T.moments(cpu["tally"], cpu_histories)
                checks = {}
                for k, image in enumerate((same.mean, same.primary_mean, same.scatter_mean)):
                    g = image.cpu().numpy().astype(np.float64).ravel()
                    c = mean[k].ravel()
                    checks[("total", "primary", "scatter")[k]] = float(np.linalg.norm(g - c) / np.linalg.norm(c))
                out[name]["tallies_vs_oracle"] = {
                    "histories": cpu_histories,
                    "relative_l2": checks,
                    "status_histogram_equal": list(same.status_histogram.counts) == cpu["histogram"].tolist(),
                    "passed": max(checks.values()) <= 1e-4,
                }
            out[name]["cpu_baseline"] = {
                "value": round(cpu_histories / elapsed, 1),
                "unit": "histories/s",
                "cores": T.threads(),
                "kind": "port",
                "sample": f"CPU oracle walk (OpenMP over histories), first {cpu_histories} histories of the same "
                          f"seed, tallies included: {elapsed:.2f} s",
            }
        _log(f"transport {name}: {out[name]['histories_per_s']:.4g} histories/s "
             f"({out[name]['wall_ms']:.1f} ms for {histories})")
        ws.close()
    return out


def precision_benchmark(grid, geometry, values, target, samples_list=(512, 1024)) -> dict:
    """Precision form of the full scan: BF16 and FP16 field storage (FP32
    sampling, compensated along-ray sums) against the FP64 path at the same
    field and pose; FP32 storage alongside.  The cotangent is the FP64
    residual against the seeded target (the squared-error loss of the full
    scan), the same for every form.  Relative L2 errors of the image, the
    volume gradient and the 6-DoF pose gradient, and the forward + joint VJP
    iteration rate."""
    import numpy as np
    import torch

    from dpt.torch import Projector

    device = torch.device("cuda:0")
    field = torch.tensor(values, device=device, requires_grad=True)
    pose = torch.zeros(6, dtype=torch.float64, device=device, requires_grad=True)
    views, (rows, cols) = len(geometry.views), geometry.shape_rows_columns

    def rel(a, b):
        a = a.detach().double().reshape(-1)
        b = b.detach().double().reshape(-1)
        return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))

    out = {}
    for samples in samples_list:
        ref_module = Projector(grid=grid, geometry=geometry, precision="fp64", samples_per_ray=samples)
        fp64_times = []
        for _ in range(2):  # the first call also creates the plan
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            image64 = ref_module(field, pose)
            cotangent = (image64 - target.double()).detach()
            g64, p64 = torch.autograd.grad(image64, (field, pose), cotangent)
            torch.cuda.synchronize()
            fp64_times.append(time.perf_counter() - t0)
        fp64_s = fp64_times[-1]
        row = {"float64_reference": {
            "iteration_ms": round(fp64_s * 1e3, 3),
            "gsamples_per_s": round(views * rows * cols * samples / fp64_s / 1e9, 3),
            "first_call_ms": round(fp64_times[0] * 1e3, 3),
            "note": "FP64 sampling of the FP32 field, forward + joint VJP, second call",
        }}
        for storage in ("float32", "bfloat16", "float16"):
            module = Projector(grid=grid, geometry=geometry, precision="fp32", samples_per_ray=samples,
                               storage=storage)
            seed32 = cotangent.float()
            times = []
            for _ in range(3):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                image = module(field, pose)
                g, p = torch.autograd.grad(image, (field, pose), seed32)
                torch.cuda.synchronize()
                times.append(time.perf_counter() - t0)
            median = sorted(times)[1]
            row[storage] = {
                "image_rel_l2": rel(image, image64),
                "volume_grad_rel_l2": rel(g, g64),
                "pose_grad_rel_l2": rel(p, p64),
                "iteration_ms": round(median * 1e3, 3),
                "gsamples_per_s": round(views * rows * cols * samples / median / 1e9, 3),
            }
            _log(f"precision S={samples} {storage}: image {row[storage]['image_rel_l2']:.2e}, volume "
                 f"{row[storage]['volume_grad_rel_l2']:.2e}, pose {row[storage]['pose_grad_rel_l2']:.2e}, "
                 f"{row[storage]['gsamples_per_s']} Gsamples/s")
        out[f"samples_{samples}"] = row
        del ref_module, image64, g64, p64
    return out


def gather_benchmark(grid, geometry, values, target, samples: int) -> dict:
    """Deterministic FP64 adjoint gather of the full scan: the cotangent is the
    FP64 residual against the seeded target; the gather (one thread per
    voxel, fixed-order FP64 sums) is timed on its second call (the first
    allocates the plan's ray table), repeated once to check bit-identity,
    and compared with the atomic FP64 volume VJP of the same cotangent."""
    import torch

    from dpt.projection import prepare
    from dpt.torch import Projector

    device = torch.device("cuda:0")
    module = Projector(grid=grid, geometry=geometry, precision="fp64", samples_per_ray=samples)
    pose = torch.zeros(6, dtype=torch.float64, device=device)
    field = torch.tensor(values, device=device)
    with torch.no_grad():
        cotangent = (module(field, pose) - target.double()).contiguous()
    entry = prepare(module._configuration, 0, torch.cuda.current_stream(device).cuda_stream, 0)
    entry.plan.set_timing(True)
    times, kernel = [], []
    outputs = []
    for _ in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        outputs.append(module.adjoint_gather(cotangent, pose))
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        kernel.append(entry.plan.kernel_ms(2))
    repeat = bool(torch.equal(outputs[0], outputs[1]))
    leaf = field.clone().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    (atomic,) = torch.autograd.grad(module(leaf, pose), leaf, cotangent)
    torch.cuda.synchronize()
    atomic_s = time.perf_counter() - t0
    a, b = outputs[1].reshape(-1), atomic.double().reshape(-1)
    rel = float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(a))
    voxels, views = grid.voxels, len(geometry.views)
    result = {
        "ms": round(times[1] * 1e3, 3),
        "kernel_ms": round(kernel[1], 3),
        "first_call_ms": round(times[0] * 1e3, 3),
        "voxel_views_per_s": round(voxels * views / times[1], 1),
        "bitwise_repeat": repeat,
        "rel_l2_vs_atomic_fp64_volume_vjp": rel,
        "atomic_fp64_forward_volume_vjp_ms": round(atomic_s * 1e3, 3),
        "output": "float64 (nz, ny, nx)",
        "note": "FP64 VALU-issue bound (one thread per voxel, culled pixel and sample windows); "
                "ms includes the per-call ray preparation",
    }
    _log(f"gather: {result['ms']} ms, repeat bitwise {repeat}, rel vs atomic {rel:.2e}")
    del outputs, atomic
    return result


# The reference's own gather workload (benchmarks/gather.py:49-160 of the
# reference, results docs/assets/results/extra-gather-chest-512-valid02.json):
# a 512 x 512 x 139 chest grid (0.762 x 0.762 x 2.5 mm), its AP view on a
# 512^2 flat panel at 2.674 mm pitch, 1024 FP64 midpoint samples, identity
# pose, and two unit-sum cotangents; device medians on the reference's GPU.
CHEST_GRID = dict(shape=(139, 512, 512), spacing_mm=(0.7617189884185791, 0.7617189884185791, 2.5),
                  origin_mm=(195.0, 171.6999969482422, -347.75),
                  orientation=(-1.0, 0.0, 0.0, 0.0, -1.0, 0.0, 0.0, 0.0, 1.0))
CHEST_AP = dict(source_mm=(0.3807984590531992, 1280.854457404986, -175.25),
                origin_mm=(683.5250937177063, -674.8060355915503, -858.3942952586534),
                u=(-1.0, 1.2246467991473532e-16, 0.0), v=(0.0, 0.0, 1.0),
                spacing_mm=(2.673754580268702, 2.673754580268702), shape=(512, 512))
CHEST_REFERENCE_MS = {"detector_region": 101.58100509643555, "full_detector_mean": 173.40975952148438}
CHEST_REFERENCE_SCATTER_MS = {"detector_region": 3.5854079723358154, "full_detector_mean": 5.735919952392578}


def gather_chest_benchmark(repeats: int = 10) -> dict:
    """Deterministic gather on the reference's chest workload.  The gather
    depends only on the geometry and the cotangent, so the field (a CT in
    the reference's run) does not enter.  Patterns: every pixel 1/N
    ("full_detector_mean") and the central 6/16 x 6/16 window normalised to
    unit sum ("detector_region").  Device time from HIP events on the launch
    stream, median of ``repeats`` calls after one warm call.  The ray-driven
    FP64 volume VJP of the same cotangent (the reference's "scatter") is
    timed beside it with torch events around the backward call: accumulator
    clear, VJP kernel and the FP64-to-FP32 gradient conversion."""
    import numpy as np
    import torch

    from dpt.geometry import DetectorGeometry
    from dpt.projection import prepare
    from dpt.torch import Projector
    from dpt.volumes import GridSpec

    device = torch.device("cuda:0")
    grid = GridSpec(**CHEST_GRID)
    geometry = DetectorGeometry(**CHEST_AP)
    module = Projector(grid=grid, geometry=geometry, precision="fp64", samples_per_ray=1024)
    pose = torch.zeros(6, dtype=torch.float64, device=device)
    rows, cols = CHEST_AP["shape"]
    result = {"grid": "512 x 512 x 139 (0.762 x 0.762 x 2.5 mm)", "detector": "512 x 512, 2.674 mm, AP view",
              "samples": 1024, "output": "float64"}
    for pattern in ("detector_region", "full_detector_mean"):
        weights = torch.ones((1, rows, cols), dtype=torch.float64, device=device)
        if pattern == "detector_region":
            weights.zero_()
            weights[:, rows * 5 // 16: rows * 11 // 16, cols * 5 // 16: cols * 11 // 16] = 1.0
        weights /= weights.sum()
        out = module.adjoint_gather(weights, pose)  # warm call: allocates the plan's ray table
        entry = prepare(module._configuration, 0, torch.cuda.current_stream(device).cuda_stream, 0)
        entry.plan.set_timing(True)
        first = out.clone()
        kernel = []
        for _ in range(repeats):
            out = module.adjoint_gather(weights, pose)
            torch.cuda.synchronize()
            kernel.append(entry.plan.kernel_ms(2))
        ms = float(np.median(kernel))
        # the ray-driven FP64 volume VJP of the same cotangent (linear in the
        # field, so a zero field serves)
        leaf = torch.zeros(CHEST_GRID["shape"], dtype=torch.float32, device=device, requires_grad=True)
        (atomic,) = torch.autograd.grad(module(leaf, pose), leaf, weights)
        scatter = []
        for _ in range(repeats):
            image = module(leaf, pose)
            start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            start.record()
            (atomic,) = torch.autograd.grad(image, leaf, weights)
            stop.record()
            torch.cuda.synchronize()
            scatter.append(start.elapsed_time(stop))
        a, b = out.reshape(-1), atomic.double().reshape(-1)
        rel = float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
        result[pattern] = {"device_median_ms": round(ms, 3), "bitwise_repeat": bool(torch.equal(first, out)),
                           "rel_l2_vs_atomic_fp64_volume_vjp": rel,
                           "atomic_volume_vjp_ms": round(float(np.median(scatter)), 3),
                           "reference_scatter_device_median_ms": CHEST_REFERENCE_SCATTER_MS[pattern],
                           "nonzero_voxels": int(torch.count_nonzero(out)),
                           "reference_device_median_ms": CHEST_REFERENCE_MS[pattern],
                           "speedup_vs_reference": round(CHEST_REFERENCE_MS[pattern] / ms, 2)}
        _log(f"chest gather {pattern}: {ms:.3f} ms (reference {CHEST_REFERENCE_MS[pattern]:.1f} ms), "
             f"rel vs atomic {rel:.2e}")
        del out, first, atomic, leaf
    return result


def cell_gauss_benchmark(grid, geometry, values, target, samples: int) -> dict:
    """Exact cell-Gauss integration of the full scan (FP64 sampling, two Gauss
    nodes per interpolation cell): forward, forward + pose VJP and forward +
    joint VJP (second calls), and the quadrature error of the FP64 midpoint
    rule at ``samples`` against it (relative L2 of the images)."""
    import torch

    from dpt.torch import Projector

    device = torch.device("cuda:0")
    exact = Projector(grid=grid, geometry=geometry, precision="fp64", samples_per_ray=samples,
                      integration="cell_gauss")
    midpoint = Projector(grid=grid, geometry=geometry, precision="fp64", samples_per_ray=samples)
    field = torch.tensor(values, device=device)
    leaf = field.clone().requires_grad_(True)
    pose = torch.zeros(6, dtype=torch.float64, device=device, requires_grad=True)

    def timed(fn, repeats=2):
        out, ms = None, []
        for _ in range(repeats):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            out = fn()
            torch.cuda.synchronize()
            ms.append((time.perf_counter() - t0) * 1e3)
        return out, ms[-1]

    with torch.no_grad():
        image, forward_ms = timed(lambda: exact(field, pose.detach()))
        reference = midpoint(field, pose.detach())
    cotangent = (image - target.double()).contiguous()
    _, pose_ms = timed(lambda: torch.autograd.grad(exact(field, pose), pose, cotangent))
    _, joint_ms = timed(lambda: torch.autograd.grad(exact(leaf, pose), (leaf, pose), cotangent))
    quadrature = float(torch.linalg.vector_norm(reference - image) / torch.linalg.vector_norm(image))
    rays = len(geometry.views) * geometry.shape_rows_columns[0] * geometry.shape_rows_columns[1]
    result = {
        "forward_ms": round(forward_ms, 3),
        "forward_pose_vjp_ms": round(pose_ms, 3),
        "forward_joint_vjp_ms": round(joint_ms, 3),
        "rays_per_s_forward": round(rays / forward_ms * 1e3, 1),
        "midpoint_quadrature_rel_l2": quadrature,
        "midpoint_samples": samples,
        "note": "FP64 exact rule; the volume VJP walks lockstep chunks of the rays, adds each piece's 8 "
                "FP32-rounded corner increments to the chunk's FP64 LDS brick and drains it with one FP64 "
                "global atomic per touched grid sample and chunk",
    }
    _log(f"cell-gauss: forward {result['forward_ms']} ms, joint {result['forward_joint_vjp_ms']} ms, "
         f"midpoint quadrature error {quadrature:.2e}")
    return result


# The reference's full-scan acceptance guards (benchmarks/config/fullscan.json
# "quality", benchmarks/config/fullscan_guards.json, benchmarks/fullscan_guards.py):
# thresholds, views, pixels, tangents, directions and steps as published there.
GUARD = {
    "image_fp32_vs_fp64": {"max_abs": 5e-4, "rmse": 2e-5, "relative_l2": 1e-5},
    "volume_adjoint_max": 2e-5,
    "views": (0, 37, 113),
    "pixels": ((127, 193), (187, 251), (256, 256), (331, 301), (401, 278)),
    "anchor_tangent": (0.137, -0.219, 0.083, 0.0017, -0.0023, 0.0009),
    "pose_direction": (0.31, -0.27, 0.19, 0.0007, -0.0011, 0.0009),
    "pose_fd": {"fp32": {"rtol": 0.02, "atol": 2e-5, "steps": (0.02, 0.04, 0.01)},
                "fp64": {"rtol": 0.005, "atol": 1e-7, "steps": (1e-6, 1e-5, 1e-7)}},
    "volume_fd": {"fp32": {"rtol": 0.005, "atol": 1e-6}, "fp64": {"rtol": 2e-5, "atol": 1e-8}},
    "volume_step": 0.01,
    "ray_tolerance": {"fp32": 5e-4, "fp64": 2e-10},
}


def guard_cotangent(views: int, rows: int, cols: int):
    """The full scan's dense objective (fullscan.json "cotangent"): raw = 1 +
    0.20 sin(theta) + 0.10 cos(2 theta) + 0.15 r + 0.10 c + 0.05 sin(theta) c,
    theta = 2 pi k / views, r and c pixel centres in [-1, 1]; divided by the
    global L1 sum in FP64, then cast once to FP32."""
    import numpy as np

    theta = 2.0 * np.pi * np.arange(views)[:, None, None] / views
    r = (2.0 * (np.arange(rows) + 0.5) / rows - 1.0)[None, :, None]
    c = (2.0 * (np.arange(cols) + 0.5) / cols - 1.0)[None, None, :]
    raw = 1.0 + 0.20 * np.sin(theta) + 0.10 * np.cos(2 * theta) + 0.15 * r + 0.10 * c + 0.05 * np.sin(theta) * c
    return (raw / np.abs(raw).sum()).astype(np.float32)


def independent_ray(values, grid, panel, pose, row, column, samples) -> float:
    """One ray on the CPU, written apart from the product path: FP64 world ->
    object -> grid transform, slab clip of the segment to the voxel-face
    support, ``samples`` midpoints, explicit eight-corner trilinear weights
    with clamped indices (the algorithm of benchmarks/fullscan_guards.py:15-53)."""
    import numpy as np

    source = np.asarray(panel.source_mm, dtype=np.float64)
    target = (np.asarray(panel.origin_mm) + column * panel.spacing_mm[0] * np.asarray(panel.u)
              + row * panel.spacing_mm[1] * np.asarray(panel.v))
    rotation = np.asarray(pose.rotation, dtype=np.float64).reshape(3, 3)
    axes = np.asarray(grid.orientation, dtype=np.float64).reshape(3, 3)
    ends = (np.stack((source, target)) - np.asarray(pose.translation_mm)) @ rotation
    ends = ((ends - np.asarray(grid.origin_mm)) @ axes) / np.asarray(grid.spacing_mm)
    start, delta = ends[0], ends[1] - ends[0]
    shape = np.asarray(values.shape[::-1])
    enter, leave = 0.0, 1.0
    for axis in range(3):
        low, high = -0.5, shape[axis] - 0.5
        if delta[axis] == 0:
            if not low <= start[axis] <= high:
                return 0.0
            continue
        a, b = sorted(((low - start[axis]) / delta[axis], (high - start[axis]) / delta[axis]))
        enter, leave = max(enter, a), min(leave, b)
    if leave <= enter:
        return 0.0
    t = enter + (np.arange(samples) + 0.5) * (leave - enter) / samples
    points = np.clip(start + t[:, None] * delta, 0, shape - 1)
    lower = np.floor(points).astype(int)
    upper = np.minimum(lower + 1, shape - 1)
    frac = points - lower
    total = np.zeros(samples)
    for bits in np.ndindex(2, 2, 2):
        b = np.asarray(bits[::-1])  # (x, y, z) corner bits
        idx = np.where(b, upper, lower)
        weight = np.prod(np.where(b, frac, 1.0 - frac), axis=1)
        total += weight * values[idx[:, 2], idx[:, 1], idx[:, 0]]
    return float(total.mean() * np.linalg.norm(target - source) * (leave - enter))


def fullscan_guards(grid, geometry, values, samples: int) -> dict:
    """The reference's acceptance guards on the full-scan workload: the FP32
    image against the FP64 image of every view, the volume adjoint identity
    <A mu, c> = <mu, A^T c> over every view with the dense objective, central
    pose and volume differences at the published steps (guard views, anchor
    pose, FP32 and FP64), and independent CPU rays on the guard pixels.  Each
    check records pass / fail against the published thresholds."""
    import numpy as np
    import torch

    from dpt.geometry import BatchedGeometry, RigidTransform, compose_pose
    from dpt.torch import Projector

    device = torch.device("cuda:0")
    views, (rows, cols) = len(geometry.views), geometry.shape_rows_columns
    out = {}
    field = torch.tensor(values, device=device, requires_grad=True)
    with torch.no_grad():
        image32 = Projector(grid=grid, geometry=geometry, samples_per_ray=samples)(field.detach())
        image64 = Projector(grid=grid, geometry=geometry, precision="fp64", samples_per_ray=samples)(field.detach())
    diff = image32.double() - image64
    bar = GUARD["image_fp32_vs_fp64"]
    check = {"max_abs": float(diff.abs().max()), "rmse": float(diff.square().mean().sqrt()),
             "relative_l2": float(torch.linalg.vector_norm(diff) / torch.linalg.vector_norm(image64))}
    check["passed"] = all(check[k] <= bar[k] for k in bar)
    out["image_fp32_vs_fp64"] = {**check, "limits": bar}
    del image64, diff
    dense = guard_cotangent(views, rows, cols)
    seed = torch.tensor(dense, device=device)
    module = Projector(grid=grid, geometry=geometry, samples_per_ray=samples)
    image = module(field)
    (field_vjp,) = torch.autograd.grad(image, field, seed)
    lhs = float((image.double() * seed.double()).sum())
    rhs = float((field.detach().double() * field_vjp.double()).sum())
    discrepancy = abs(lhs - rhs) / max(abs(lhs), abs(rhs), 1e-300)
    out["volume_adjoint"] = {"lhs": lhs, "rhs": rhs, "relative_discrepancy": discrepancy,
                             "limit": GUARD["volume_adjoint_max"], "passed": discrepancy <= GUARD["volume_adjoint_max"]}
    del image, field_vjp
    picks = GUARD["views"]
    subset = BatchedGeometry(tuple(geometry.views[k] for k in picks))
    anchor = compose_pose(RigidTransform(), GUARD["anchor_tangent"])
    seed_host = dense[list(picks)].astype(np.float64)
    seed_host /= np.abs(seed_host).sum()
    direction = torch.tensor(GUARD["pose_direction"], dtype=torch.float64, device=device)
    z, y, x = np.ogrid[:grid.shape[0], :grid.shape[1], :grid.shape[2]]
    base = values.astype(np.float64)
    step = GUARD["volume_step"]
    perturbation = base * 0.01 + (base > 0) * 1e-5 * (1 + 0.2 * np.cos(x / 13) + 0.2 * np.cos(y / 17)
                                                       + 0.2 * np.cos(z / 19))
    perturbation = np.minimum(perturbation, 0.5 * base / step)
    plus_host = (base + step * perturbation).astype(np.float32)
    minus_host = (base - step * perturbation).astype(np.float32)
    realised = (plus_host.astype(np.float64) - minus_host.astype(np.float64)) / (2 * step)
    for precision in ("fp32", "fp64"):
        projector = Projector(grid=grid, geometry=subset, anchor=anchor, precision=precision,
                              samples_per_ray=samples)
        dtype = torch.float32 if precision == "fp32" else torch.float64
        weights = torch.tensor(seed_host, dtype=dtype, device=device)
        tangent = torch.zeros(6, dtype=torch.float64, device=device, requires_grad=True)
        image = projector(field, tangent)
        volume_grad, pose_grad = torch.autograd.grad(image, (field, tangent), weights)

        def objective(output):
            return float((output.double() * weights.double()).sum())

        limits = GUARD["pose_fd"][precision]
        analytic = float((pose_grad * direction).sum())
        checks = []
        with torch.no_grad():
            for h in limits["steps"]:
                fd = (objective(projector(field.detach(), h * direction))
                      - objective(projector(field.detach(), -h * direction))) / (2 * h)
                error = abs(analytic - fd)
                allowed = limits["atol"] + limits["rtol"] * abs(analytic)
                checks.append({"step": h, "analytic": analytic, "finite_difference": fd, "absolute_error": error,
                               "allowed_absolute_error": allowed, "passed": error <= allowed,
                               "acceptance_step": h == limits["steps"][0]})
        out[f"pose_fd_{precision}"] = {"checks": checks, "passed": checks[0]["passed"]}
        bound = GUARD["volume_fd"][precision]
        with torch.no_grad():
            plus = objective(projector(torch.tensor(plus_host, device=device), tangent.detach()))
            minus = objective(projector(torch.tensor(minus_host, device=device), tangent.detach()))
        analytic_v = float(np.sum(volume_grad.cpu().numpy().astype(np.float64) * realised))
        fd_v = (plus - minus) / (2 * step)
        error = abs(analytic_v - fd_v)
        allowed = bound["atol"] + bound["rtol"] * abs(analytic_v)
        out[f"volume_fd_{precision}"] = {"analytic": analytic_v, "finite_difference": fd_v, "absolute_error": error,
                                         "allowed_absolute_error": allowed, "passed": error <= allowed, "step": step}
        actual = image.detach().cpu().numpy()
        tolerance = GUARD["ray_tolerance"][precision]
        rays = []
        for local, view in enumerate(picks):
            for row, column in GUARD["pixels"]:
                expected = independent_ray(values, grid, subset.views[local], anchor, row, column, samples)
                observed = float(actual[local, row, column])
                rays.append({"view": view, "pixel": [row, column], "reference": expected, "actual": observed,
                             "absolute_error": abs(observed - expected), "passed": abs(observed - expected) <= tolerance})
        out[f"independent_rays_{precision}"] = {"tolerance": tolerance, "rays": rays,
                                               "max_absolute_error": max(r["absolute_error"] for r in rays),
                                               "passed": all(r["passed"] for r in rays)}
        del projector, image, volume_grad, pose_grad
    out["passed"] = all(v["passed"] for v in out.values() if isinstance(v, dict) and "passed" in v)
    _log("guards: " + ", ".join(f"{k} {'pass' if v['passed'] else 'FAIL'}" for k, v in out.items()
                                if isinstance(v, dict)))
    return out


def roofline_report(args, fwd_ms: float, vjp_ms: float, effective: int, half: bool) -> dict:
    """Roofline of the dominant kernel (the joint VJP at the benchmark's
    shape).  ``achieved`` is its measured HBM traffic per launch (rocprofv3
    FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel, recorded in
    profiles/projector_counters.json for the sources digest it was measured
    at) over its launch time (HIP events on its launch stream in the timed
    region); no figure is given when the recorded digest is not the benched
    sources'.  The gathers and scatters' algorithmic bytes (64 B per effective
    sample, mostly served from L1/L2/LDS) are reported as a rate beside it,
    not as a fraction.  The kernel's binding resource is instruction issue:
    ``valu`` sets its VALU instruction rate (counted instructions over launch
    time) against the issue rate of independent v_fma_f32 / v_add_u32 chains
    measured on this GPU in this run."""
    import torch

    from dpt import _native

    gather_b = 16 if half else GATHER_B
    if vjp_ms >= fwd_ms:
        kernel, ms, per_sample = "vjp32_kernel<float, true, true, 12, false, 2, false>", vjp_ms, gather_b + SCATTER_B
    else:
        kernel, ms, per_sample = "forward32_kernel<float, false, 2, false>", fwd_ms, gather_b
    digest = sources_digest()
    recorded, counters = None, None
    if COUNTERS.exists():
        recorded = json.loads(COUNTERS.read_text())
        if recorded.get("sources_sha256") == digest and not half and args.samples == SAMPLES and \
                args.views == VIEWS:
            counters = next((v for k, v in recorded["kernels"].items() if kernel in k), None)
    seconds = ms * 1e-3
    report = {"bound": "hbm", "kernel": kernel, "kernel_ms": round(ms, 3), "forward_kernel_ms": round(fwd_ms, 3),
              "vjp_kernel_ms": round(vjp_ms, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
              "achieved": None, "frac": None, "traffic": None,
              "counters": {"profile": str(COUNTERS.relative_to(ROOT)), "sources_sha256": digest,
                           "recorded_sha256": recorded.get("sources_sha256") if recorded else None,
                           "matches": counters is not None}}
    if counters is not None and "fetch_bytes" in counters and "write_bytes" in counters:
        traffic = counters["fetch_bytes"] + counters["write_bytes"]
        report["traffic"] = round(traffic)
        report["achieved"] = round(traffic / seconds / 1e9, 1)
        report["frac"] = round(traffic / seconds / 1e9 / HBM_PEAK_GBS, 4)
    report["algorithmic_bytes_per_effective_sample"] = per_sample
    report["algorithmic_rate_gbs"] = round(effective * per_sample / seconds / 1e9, 1)
    scratch = torch.empty(2048 * 256, dtype=torch.float32, device="cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    fma = _native.valu_probe(0, scratch.data_ptr(), stream)
    add = _native.valu_probe(1, scratch.data_ptr(), stream)
    valu = {"unit": "wave64 VALU instructions/s", "peak": round(max(fma, add), 1),
            "probe_v_fma_f32": round(fma, 1), "probe_v_add_u32": round(add, 1), "achieved": None, "frac": None}
    if counters is not None and "SQ_INSTS_VALU" in counters:
        rate = counters["SQ_INSTS_VALU"] / seconds
        valu["achieved"] = round(rate, 1)
        valu["frac"] = round(rate / valu["peak"], 4)
        valu["lane_instructions_per_effective_sample"] = round(counters["SQ_INSTS_VALU"] * 64 / effective, 1)
    report["valu"] = valu
    return report


def _log(message: str) -> None:
    print(f"[bench] {message}", file=sys.stderr, flush=True)


def gpu_parity(grid, geometry, values, samples, picks, image, seed, grad_field, grad_pose) -> dict:
    """The HIP path (default FP32 plan, through the C ABI) on the same views,
    field and cotangent as the CPU oracle run: relative L2 of the image and of
    the volume and pose gradients (the correctness gate of the timed
    workload, at its full detector size and sample count)."""
    import numpy as np
    import torch

    from dpt import _native

    panels = np.asarray([geometry.views[k].packed() for k in picks], dtype=np.float64)
    rows, cols = geometry.shape_rows_columns
    plan = _native.Plan(grid, views=len(picks), rows=rows, cols=cols, panels=panels, rays=None, samples=samples,
                        arithmetic="float32", storage="float32",
                        needs=_native.VOLUME | _native.POSE | _native.GATHER_CELLS, device=0)
    stream = torch.cuda.current_stream().cuda_stream
    field = torch.tensor(values, device="cuda:0")
    poses = torch.tensor([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], dtype=torch.float64,
                         device="cuda:0").expand(len(picks), 12).contiguous()
    out = torch.empty((len(picks), rows, cols), device="cuda:0")
    plan.forward(field.data_ptr(), poses.data_ptr(), out.data_ptr(), stream)
    cot = torch.tensor(np.ascontiguousarray(seed, dtype=np.float32), device="cuda:0")
    gv = torch.empty(grid.shape, device="cuda:0")
    gp = torch.empty((len(picks), 12), dtype=torch.float64, device="cuda:0")
    plan.vjp(field.data_ptr(), poses.data_ptr(), cot.data_ptr(), gv.data_ptr(), gp.data_ptr(),
             _native.VOLUME | _native.POSE, stream)
    status = plan.status(stream)
    plan.close()

    def rel(a, b):
        a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    gpu_image = out.cpu().numpy()
    return {
        "views": picks,
        "image_rel_l2": rel(gpu_image, image),
        "image_bitwise_equal_fraction": float(np.mean(gpu_image == image)),
        "volume_grad_rel_l2": rel(gv.cpu().numpy(), grad_field),
        "pose_grad_rel_l2": rel(gp.cpu().numpy(), grad_pose),
        "status": [int(c) for c in status],
        "bar": "1e-5 relative L2 (FP32 path)",
    }


def cpu_baseline(grid, geometry, values, samples: int, stride: int) -> dict:
    """Time the CPU oracle (FP32 restatement) on views {0,90,180,270}, every
    ``stride``-th detector row, full columns and samples: forward + joint VJP,
    on all host threads; plus the same per-sample work on one thread (view 0,
    every 16th row) and the small CPU-runnable case (64^3, AP + lateral,
    128^2 at 0.8 mm, 256 samples, central-region mean loss) in full."""
    import math

    import numpy as np

    from dpt import phantoms
    from dpt.geometry import ConeBeamIntrinsics
    from oracle import oracle as O

    def run(picks, row_stride, scene_grid, scene_values, geom, s, seed_fn):
        panels = []
        for k in picks:
            p = geom.views[k]
            rows = (p.shape[0] + row_stride - 1) // row_stride
            panels.append(O.Panel(p.source_mm, p.origin_mm, p.u, p.v, (p.spacing_mm[0], p.spacing_mm[1] * row_stride),
                                  (rows, p.shape[1])))
        scene = O.Scene(O.Grid(scene_grid.shape, scene_grid.spacing_mm, scene_grid.origin_mm,
                               scene_grid.orientation), s, panels=panels)
        t0 = time.perf_counter()
        image, _ = scene.forward(scene_values)
        seed = seed_fn(image)
        g, pg, _, _ = scene.vjp(scene_values, seed)
        elapsed = time.perf_counter() - t0
        nominal = len(picks) * panels[0].shape[0] * panels[0].shape[1] * s
        run.last = (image, seed, g, pg)
        return nominal, elapsed, panels[0].shape

    residual_like = lambda image: (image - 0.5 * image.mean()).astype(np.float32)  # noqa: E731
    picks = [v for v in (0, 90, 180, 270) if v < len(geometry.views)]
    nominal, elapsed, shape = run(picks, stride, grid, values, geometry, samples, residual_like)
    parity = gpu_parity(grid, geometry, values, samples, picks, *run.last) if stride == 1 else None
    threads = O.threads()
    O.set_threads(1)
    try:
        one_nominal, one_elapsed, one_shape = run([0], 16, grid, values, geometry, samples, residual_like)
    finally:
        O.set_threads(threads)
    # CPU-runnable case: 64^3, AP and lateral, 128^2 at 0.8 mm, S = 256,
    # loss = mean of the central 64x64 of each view
    small = phantoms.centred_grid((64, 64, 64))
    small_values = phantoms.chest_phantom(small, seed=2026)
    angles = phantoms.standard_angles()
    small_geometry = ConeBeamIntrinsics(SOD, SDD, (128, 128), (0.8, 0.8), centre_mm=small.centre_mm()).batch(
        [angles["ap"], angles["lateral"]])

    def central_mean(image):
        seed = np.zeros(image.shape, dtype=np.float32)
        seed[:, 32:96, 32:96] = 1.0 / (image.shape[0] * 64 * 64)
        return seed

    c1_nominal, c1_elapsed, _ = run([0, 1], 1, small, small_values, small_geometry, 256, central_mean)
    return {
        "value": round(nominal / elapsed / 1e9, 6),
        "unit": "Gsamples/s",
        "cores": threads,
        "kind": "port",
        "sample": (
            f"CPU oracle (reference-algorithm FP32 restatement, OpenMP over rays) forward + volume + pose VJP on views "
            f"{picks}, every {stride}th detector row ({shape[0]}x{shape[1]} rays/view), "
            f"{samples} samples/ray: {nominal:.3e} nominal samples in {elapsed:.2f} s; a full 360-view pass "
            f"extrapolates to {elapsed * len(geometry.views) / len(picks) * stride:.0f} s"
        ),
        **({"parity": parity} if parity else {}),
        "single_thread": {
            "value": round(one_nominal / one_elapsed / 1e9, 6),
            "unit": "Gsamples/s",
            "cores": 1,
            "sample": f"view 0, every 16th row ({one_shape[0]}x{one_shape[1]} rays), {samples} samples/ray: "
                      f"{one_nominal:.3e} samples in {one_elapsed:.2f} s",
        },
        "small_case": {
            "value": round(c1_nominal / c1_elapsed / 1e9, 6),
            "unit": "Gsamples/s",
            "cores": threads,
            "iteration_ms": round(c1_elapsed * 1e3, 3),
            "sample": "64^3 phantom, AP + lateral, 128x128 at 0.8 mm, 256 samples, central 64x64 mean loss, in full",
        },
    }


def main() -> int:
    args = _args()
    if args.gpus != 1:
        print(
            json.dumps({"error": "this build measures one GPU; sharding across GPUs is described, not run",
                        "n_gpus": args.gpus}),
            flush=True,
        )
        return 2
    if args.record_counters:
        return record_counters(args.record_counters, args.calibration_bytes)
    if args.copy_calibration > 0:
        return copy_calibration(args.copy_calibration)
    if args.request_only:
        print(json.dumps(request_benchmark(args.request_repeats)), flush=True)
        return 0
    if args.serving_only:
        rates = [float(r) for r in args.serving_rates.split(",") if r]
        forms = tuple(f for f in args.serving_forms.split(",") if f)
        open_forms = tuple(f for f in forms if f in ("batched_bf16", "plain_fp32"))
        print(json.dumps(serving_benchmark(args.serving_requests, rates, forms=forms, open_forms=open_forms)),
              flush=True)
        return 0
    if args.kernels_once:
        return kernels_once(args)
    if args.transport_only:
        print(json.dumps(transport_benchmark(args.transport_histories, args.transport_cpu_histories,
                                             not args.no_cpu_baseline)), flush=True)
        return 0
    if args.gather_chest_only:
        print(json.dumps(gather_chest_benchmark()), flush=True)
        return 0
    import numpy as np
    import torch

    from dpt import phantoms
    from dpt.projection import decode, gather_cells, prepare
    from dpt.torch import Projector

    device = torch.device("cuda:0")
    _log("building phantom and geometry")
    grid = phantoms.centred_grid(SHAPE)
    values = phantoms.chest_phantom(grid, seed=2026)
    geometry, pitch = phantoms.circular_geometry(grid, args.views, DETECTOR, sod=SOD, sdd=SDD)
    field = torch.tensor(values, device=device, requires_grad=True)

    def make():
        return Projector(grid=grid, geometry=geometry, precision="fp32", samples_per_ray=args.samples, storage=args.storage)

    # seeded target: the same phantom seen at a seeded pose (+/-2 deg, +/-5 mm) plus noise
    rng = np.random.default_rng(7)
    xi_true = np.concatenate([rng.uniform(-5, 5, 3), np.deg2rad(rng.uniform(-2, 2, 3))])
    with torch.no_grad():
        target = make()(field.detach(), torch.tensor(xi_true, dtype=torch.float64, device=device))
        target += 1e-3 * torch.randn(target.shape, generator=torch.Generator(device).manual_seed(11), device=device)
    torch.cuda.synchronize()
    if args.gather_only:
        print(json.dumps(gather_benchmark(grid, geometry, values, target, args.samples)), flush=True)
        return 0
    if args.cell_gauss_only:
        print(json.dumps(cell_gauss_benchmark(grid, geometry, values, target, args.samples)), flush=True)
        return 0

    pose = torch.zeros(6, dtype=torch.float64, device=device, requires_grad=True)
    projector = make()
    needs = 3
    stream = torch.cuda.current_stream(device).cuda_stream
    entry_fwd = prepare(projector._configuration, 0, stream, needs)
    entry_fwd.plan.set_timing(True)
    poses_packed = torch.tensor([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], dtype=torch.float64, device=device).expand(args.views, 12).contiguous()
    hits = entry_fwd.plan.count_hit_rays(poses_packed.data_ptr(), stream)
    nominal = args.views * DETECTOR[0] * DETECTOR[1] * args.samples
    effective = hits * args.samples
    _log(f"pitch {pitch:.5f} mm, hit rays {hits} of {args.views * DETECTOR[0] * DETECTOR[1]}")

    def step():
        module = make()
        image = module(field, pose)
        residual = image - target
        return torch.autograd.grad(image, (field, pose), residual)

    for i in range(args.warmup):
        step()
        torch.cuda.synchronize()
        _log(f"warmup {i + 1}/{args.warmup} done")

    times, fwd_ms, vjp_ms = [], [], []
    # Python's cyclic collector runs between steps, not inside them
    gc.collect()
    gc.disable()
    torch.cuda.synchronize()
    start = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        grads = step()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        fwd_ms.append(entry_fwd.plan.kernel_ms(0))
        vjp_ms.append(entry_fwd.plan.kernel_ms(1))
        _log(f"step {i + 1}/{args.steps}: {times[-1] * 1e3:.1f} ms (forward kernel {fwd_ms[-1]:.1f} ms, vjp kernel {vjp_ms[-1]:.1f} ms)")
    total = time.perf_counter() - start
    gc.enable()
    assert all(torch.isfinite(g).all() for g in grads)

    # the reference's component table (docs/performance/fullscan.md:248-255,
    # 323): completed resident calls, median of 3, the same Projector
    def timed(fn):
        ts = []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return sorted(ts)[1]

    module = make()
    frozen_field, frozen_pose = field.detach(), pose.detach()
    with torch.no_grad():
        seed_image = module(frozen_field, frozen_pose) - target

    def forward_only():
        with torch.no_grad():
            module(frozen_field, frozen_pose)

    def volume_vjp():
        image = module(field, frozen_pose)
        torch.autograd.grad(image, field, seed_image)

    def pose_vjp():
        image = module(frozen_field, pose)
        torch.autograd.grad(image, pose, seed_image)

    def ray_vjp():
        module.ray_vjp(frozen_field, seed_image, frozen_pose)

    components = {}
    for name, fn, ref_ms in (("forward", forward_only, 512.60), ("forward_volume_vjp", volume_vjp, 1058.0),
                             ("forward_pose_vjp", pose_vjp, None), ("ray_coordinate_vjp", ray_vjp, None)):
        t = timed(fn)
        components[name] = {"ms": round(t * 1e3, 3), "gsamples_per_s": round(nominal / t / 1e9, 3)}
        if ref_ms is not None:
            components[name]["reference_blackwell_ms"] = ref_ms
    _log("components: " + ", ".join(f"{k} {v['ms']} ms" for k, v in components.items()))

    median = float(np.median(times))
    gsamples = nominal / median / 1e9
    fwd = float(np.mean(fwd_ms))
    vjp = float(np.mean(vjp_ms))
    half = args.storage != "float32"
    roofline = roofline_report(args, fwd, vjp, effective, half)
    result = {
        "metric": "Gsamples/s fwd+vol+pose VJP 256^3/360v/512^2; p50/p99 batched-solve latency",
        "value": round(gsamples, 3),
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total / args.steps * 1e3, 3),
        "median_ms": round(median * 1e3, 3),
        "p95_ms": round(float(np.percentile(times, 95)) * 1e3, 3),
        "step_ms": [round(t * 1e3, 3) for t in times],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if not half else f"f32 sampling, {args.storage} storage",
        "data": "synthetic: seeded ellipsoid phantom (body/2 lungs/5 bones, noise 1e-4), seeded target pose",
        "config": {
            "workload": "FP32 forward + volume VJP + 6-DoF pose VJP, squared-error loss",
            "volume": "256^3 float32 at 1 mm",
            "views": args.views,
            "detector": "512x512",
            "pitch_mm": round(pitch, 6),
            "sod_sdd_mm": [SOD, SDD],
            "samples_per_ray": args.samples,
            "nominal_samples_per_pass": nominal,
            "effective_samples_per_pass": effective,
            "storage": args.storage,
            "gather": "cells" if gather_cells(decode(make()._configuration)) else "linear",
        },
        "roofline": roofline,
        "reference_blackwell_iteration_gsamples": 65.5,
        "components": components,
    }
    if not args.no_guards:
        _log("full-scan acceptance guards")
        result["guards"] = fullscan_guards(grid, geometry, values, args.samples)
    if not args.no_request:
        _log("single request (chest AP)")
        result["request"] = request_benchmark(args.request_repeats)
    if not args.no_serving:
        _log("serving benchmark")
        rates = [float(r) for r in args.serving_rates.split(",") if r]
        result["serving"] = serving_benchmark(args.serving_requests, rates)
        served = result["serving"]["forms"]["batched_bf16"]["open_loop"][f"{rates[0]:g}_per_s"]
        result["p50_ms"] = served["p50_latency_ms"]
        result["p99_ms"] = served["p99_latency_ms"]
        result["latency_rate_per_s"] = rates[0]
    if not args.no_precision:
        _log("precision benchmark")
        result["precision"] = precision_benchmark(grid, geometry, values, target)
    if not args.no_cell_gauss:
        _log("exact cell-Gauss integration")
        result["cell_gauss"] = cell_gauss_benchmark(grid, geometry, values, target, args.samples)
    if not args.no_gather:
        _log("deterministic adjoint gather")
        result["gather"] = gather_benchmark(grid, geometry, values, target, args.samples)
        result["gather_chest"] = gather_chest_benchmark()
    if not args.no_transport:
        _log("transport benchmark")
        result["transport"] = transport_benchmark(args.transport_histories, args.transport_cpu_histories,
                                                  not args.no_cpu_baseline)
    if not args.no_cpu_baseline:
        _log("timing CPU oracle baseline")
        result["cpu_baseline"] = cpu_baseline(grid, geometry, values, args.samples, args.cpu_rows_stride)
    result["children_at_exit"] = _children()
    print(json.dumps(result), flush=True)
    return 0


def _children() -> list:
    """Processes this run started that are still alive (expected: none)."""
    try:
        import psutil
    except ImportError:
        return []
    return [" ".join(c.cmdline()[:4]) or c.name() for c in psutil.Process().children(recursive=True)]


if __name__ == "__main__":
    sys.exit(main())
