"""Benchmark of the 2-D cross-correlation PIV hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): one FP32
2048 x 2048 frame pair at 0.05 particles/px (Gaussian sigma 1.2 px, peak
U[0.5, 1], N(0, 0.02) noise), frame 2 displaced by a Lamb-Oseen vortex (core
radius 300 px, peak 8 px) plus a (1.5, 0.5) px drift.  One step runs PIV on that
resident pair four times: ZNCC and NSQECC, each at interSize 32 and 64, with
searchMargin 16, step 16 and computeSN -- 16,129 + 16,129 + 15,625 + 15,625 =
63,508 IA correlations (vectors with subpixel peak and SN).  The four runs are
four requests on one device frame, issued as one batched call per step (the
second image's integral arrays are built once for all four, then one kernel
launch per run); "per_call" reports the same runs as four separate calls.

value = IA correlations per second over the timed steps (inputs resident in
HBM, outputs written to HBM).  Each kernel is timed with HIP events on the
stream it runs on; the roofline entry is for the kernel with the largest total
time.  The CPU baseline is the Float64 oracle's FFT mode (a restatement of the
reference's PIV_CPU loop with per-IA transforms) over the same four runs,
OpenMP over IAs on the host cores.

"transform_forms" times the four runs once more as single calls in both GPU
transform forms: FP32 radix transforms (mode "fp32") and the
split-precision matrix-core form (mode "mfma", DFT-matrix products of hi + lo
FP16 operands), with the latter's roofline against the dense FP16 matrix peak
on the dense-DFT flop count.

A batched leg ("batched_masked") runs 16 uneven masked NSQECC pairs (512^2 to
4096^2, interSize 16/32/64) as one captured HIP graph and as 16 separate calls.
A 3-D leg ("volume_3d") times FFTCC and ZNCC on FP32 512 x 512 x 256 particle
volumes (interSize 32^3, searchMargin 8, step 16) with its own CPU sample.
A serving leg follows ("serving"): a Poisson stream (200 pairs/s) of mixed 2-D
and 3-D requests over frames resident in HBM, served by the continuous-batching
server and by the plain per-request path on the same schedule, then the same
requests all at once (saturation); p50/p99 latency and throughput of both, and
a bit-for-bit comparison of their outputs.

    python bench.py [--steps K] [--warmup W] [--mode auto|fp32|fp64] [--no-cpu] [--no-serve] [--no-3d] [--no-batched]
    python bench.py --serve [--serve-requests N] [--serve-rate R]    (serving leg only)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import subprocess
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "IA correlations/sec (2D 32², 3D 32³) at stated px error; p50/p99 latency"
RUNS = [("zncc", 32), ("nsqecc", 32), ("zncc", 64), ("nsqecc", 64)]
SM, STEP, SIZE = 16, 16, 2048

# MI355X published peaks
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
PEAK_FP16_MFMA_TFLOPS = 2500.0  # dense


def flop_fft_per_ia(cs: int, ndims: int = 2, transforms: int = 3) -> float:
    """Algorithmic flops of one IA on the plain transform path at the reference
    transform size cs: T*2.5*P*log2(P) + (T/3)*(6*Pc + P)."""
    P = cs ** ndims
    h = cs // 2 + 1
    Pc = h * cs ** (ndims - 1)
    return transforms * 2.5 * P * math.log2(P) + (transforms / 3) * (6 * Pc + P)


def flop_dft_per_ia(cs: int, transforms: int = 3) -> float:
    """Algorithmic flops of one 2-D IA as dense DFT-matrix products at cs (the
    unpruned contraction; the hi/lo split does not count):
    T*[4*h*cs^2 + 8*cs^2*h] + (T/3)*(6*Pc + P), h = cs/2 + 1."""
    h = cs // 2 + 1
    P = cs * cs
    Pc = h * cs
    return transforms * (4 * h * cs * cs + 8 * cs * cs * h) + (transforms / 3) * (6 * Pc + P)


def form_of(mode: str, alg: str, isz: int, compute_sn: bool) -> str:
    """The transform form the library runs one of the headline runs in
    ("f32", "f64", "split-f16" or "split-f16+f64"; mqpiv_transform_form)."""
    from multi_quickpiv_amd import setPIVParameters
    from multi_quickpiv_amd.piv import transform_form
    p = setPIVParameters(corr_alg=alg, interSize=isz, searchMargin=SM, step=STEP, computeSN=compute_sn)
    return transform_form(p, (SIZE, SIZE), mode)[0]


class _ResidentRequest:
    """A request over frames shared in HBM (column-major), with its own outputs."""

    def __init__(self, frame, p, mode, out_sets=1):
        import torch

        from multi_quickpiv_amd import _lib
        from multi_quickpiv_amd.params import get_vectorfield_size
        from multi_quickpiv_amd.piv import to_c_params

        self.frame, self.p = frame, p
        shape = frame["shape"]
        p.ndims = len(shape)
        self.vf = get_vectorfield_size(shape, p)
        self.n = int(np.prod(self.vf))
        self.params = to_c_params(p, None, mode)
        self.outs = [(torch.zeros(self.n * len(shape), dtype=torch.float64, device="cuda"),
                      torch.zeros(self.n, dtype=torch.float64, device="cuda") if p.computeSN else None)
                     for _ in range(out_sets)]
        self._lib = _lib

    def c_request(self, k=0):
        L = self._lib
        r = L.MqpivRequest()
        r.params = self.params
        shape = self.frame["shape"]
        for d in range(3):
            r.dims[d] = shape[d] if d < len(shape) else 1
        r.in_dtype, r.out_dtype = L.MQPIV_F32, L.MQPIV_F64
        r.img1, r.img2 = self.frame["A"].data_ptr(), self.frame["B"].data_ptr()
        r.mask = r.skip_mask = r.peaks = None
        VF, SN = self.outs[k]
        r.VF = VF.data_ptr()
        r.SN = SN.data_ptr() if SN is not None else None
        return r


def serve_bench(n_requests: int, rates, seed: int, n_slots: int = 4, max_batch: int = 32,
                budget_gib: int = 64, mode: str = "auto", log=None) -> dict:
    """Poisson stream of mixed 2-D / 3-D requests over frames resident in HBM,
    served by the continuous-batching server and, on the same arrival schedule,
    by the plain per-request path (one pair at a time, in arrival order, each
    call synchronised), for each arrival rate in ``rates`` (0 = every request
    at once).  Returns latency percentiles (arrival to results in HBM) and
    throughput of both, and whether every served output equals the plain
    path's bit for bit."""
    import threading

    import torch

    from multi_quickpiv_amd import _lib, setPIVParameters, synth
    from multi_quickpiv_amd.piv import Server

    rng = np.random.default_rng(seed)
    frames = []
    # 2-D: 512^2 .. 4096^2 at 0.05 particles/px, Lamb-Oseen (core radius and
    # peak scaled with the frame) + drift; 3-D: 128^3 .. 512x512x128 at 0.01
    # particles/voxel, shear + drift.  Two frame pairs per size, seeded.
    for k, n in enumerate([512, 1024, 2048, 4096] * 2):
        sh = (n, n)
        a, b = synth.particle_pair(sh, synth.n_particles(sh, 0.05), 1.2,
                                   synth.lamb_oseen(sh, core_radius=300.0 * n / 2048, peak=8.0), 7000 + k,
                                   device="cuda", dtype=torch.float32)
        frames.append({"shape": sh, "A": a.t().contiguous(), "B": b.t().contiguous(), "nd": 2})
    for k, sh in enumerate([(128, 128, 128), (256, 256, 128), (512, 512, 128)] * 2):
        a, b = synth.particle_pair(sh, synth.n_particles(sh, 0.01), 1.5, synth.shear3d(), 8000 + k,
                                   device="cuda", dtype=torch.float32)
        frames.append({"shape": sh, "A": a.permute(2, 1, 0).contiguous(), "B": b.permute(2, 1, 0).contiguous(),
                       "nd": 3})
    del a, b
    f2 = [f for f in frames if f["nd"] == 2]
    f3 = [f for f in frames if f["nd"] == 3]
    algs = ["fft", "zncc", "nsqecc"]
    reqs = []
    for i in range(n_requests):
        if rng.random() < 0.7:
            fr = f2[rng.integers(len(f2))]
            isz = int(rng.choice([16, 32, 64]))
            p = setPIVParameters(corr_alg=algs[rng.integers(3)], interSize=isz, searchMargin=isz // 2,
                                 step=isz // 2)
        else:
            fr = f3[rng.integers(len(f3))]
            p = setPIVParameters(corr_alg=algs[rng.integers(3)], interSize=32, searchMargin=8, step=16)
        reqs.append(_ResidentRequest(fr, p, mode, out_sets=2))
    gaps = rng.exponential(1.0, n_requests)
    torch.cuda.synchronize()

    class _Sub:  # submit_device wants an object with c_request()
        def __init__(self, r, k):
            self.r, self.k = r, k

        def c_request(self):
            return self.r.c_request(self.k)

    big3 = max((r for r in reqs if len(r.frame["shape"]) == 3), key=lambda r: r.n)
    big2 = max((r for r in reqs if len(r.frame["shape"]) == 2), key=lambda r: r.n)
    warm = [big3] * n_slots + [big2] * n_slots  # admitted round-robin: one of each per slot
    sv = Server(0, budget_gib << 30, n_slots, max_batch)
    for r in warm:  # warm up every slot (3-D transform scratch, job tables)
        sv.wait(sv.submit_device(_Sub(r, 0)))
    ctx = _lib.context(0)
    s = torch.cuda.Stream()

    def plain_call(r):
        ctx.check(ctx.lib.mqpiv_piv_batch_device(ctx.handle, (_lib.MqpivRequest * 1)(r.c_request(1)), 1,
                                                 s.cuda_stream))

    for r in warm:
        plain_call(r)
    s.synchronize()
    # every request once through both paths, untimed: the first launch of each
    # kernel instance loads its code, which neither timed run should pay
    for t in [sv.submit_device(_Sub(r, 0)) for r in reqs]:
        sv.wait(t)
    for r in reqs:
        plain_call(r)
    s.synchronize()
    n_ia = sum(r.n for r in reqs)
    out = {"requests": n_requests, "ia_total": n_ia, "slots": n_slots, "max_batch": max_batch,
           "mix": "70% 2-D (512^2-4096^2, FFTCC/ZNCC/NSQECC, IA 16/32/64, margin = step = IA/2), "
                  "30% 3-D (128^3-512x512x128, IA 32^3, margin 8, step 16); frames resident in HBM "
                  "(8 2-D and 6 3-D seeded pairs), parameters drawn per request"}
    for rate in rates:
        arrivals = np.cumsum(gaps) / rate if rate > 0 else np.zeros(n_requests)
        # ---- served: a submitter thread follows the arrival schedule
        sv.reset_stats()
        tickets = [None] * n_requests
        t0 = time.perf_counter() + 0.05

        def submitter():
            for i, r in enumerate(reqs):
                dt = t0 + arrivals[i] - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                tickets[i] = sv.submit_device(_Sub(r, 0))

        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        for t in tickets:
            sv.wait(t)
        st = sv.stats()
        # ---- plain per-request path on the same schedule
        lat = []
        t0 = time.perf_counter() + 0.05
        t_end = t0
        for i, r in enumerate(reqs):
            dt = t0 + arrivals[i] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            plain_call(r)
            s.synchronize()
            t_end = time.perf_counter()
            lat.append(t_end - (t0 + arrivals[i]))
        lat = np.array(lat) * 1e3
        plain_span = t_end - t0
        same = all(torch.equal(r.outs[0][0], r.outs[1][0]) and
                   (r.outs[0][1] is None or torch.equal(r.outs[0][1], r.outs[1][1])) for r in reqs)
        key = f"rate_{int(rate)}" if rate > 0 else "burst"
        out[key] = {
            "arrivals": f"Poisson, {rate:g} pairs/s" if rate > 0 else "all at t=0 (saturation)",
            "served": {"p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
                       "max_ms": round(st["max_ms"], 3), "pairs_per_s": round(st["pairs_per_s"], 1),
                       "ia_per_s": round(st["ia_per_s"], 1)},
            "plain": {"p50_ms": round(float(np.percentile(lat, 50)), 3),
                      "p99_ms": round(float(np.percentile(lat, 99)), 3),
                      "max_ms": round(float(lat.max()), 3),
                      "pairs_per_s": round(n_requests / plain_span, 1),
                      "ia_per_s": round(n_ia / plain_span, 1)},
            "identical_outputs": bool(same),
        }
        if log:
            log(f"{key}: {out[key]}")
    sv.close()
    return out


def volume_bench(mode: str, steps: int, cpu: bool, nthreads: int) -> dict:
    """The 3-D workload: FP32 512 x 512 x 256 particle volumes at 0.01
    particles/voxel (sigma 1.5), shear u = 0.01 z plus drift (2.3, -1.1, 0.7);
    FFTCC and ZNCC at interSize 32^3, searchMargin 8, step 16 (14,415 IAs, the
    reference transform size 48^3).  IA/s over `steps` timed calls with the
    volumes resident in HBM, kernel time from the library's events, and the
    Float64 oracle (FFT mode) on a sample of the same IAs."""
    import torch

    from multi_quickpiv_amd import _lib, setPIVParameters, synth
    from multi_quickpiv_amd.params import get_vectorfield_size
    from multi_quickpiv_amd.piv import piv_device

    a, b = synth.config_c3(device="cuda")
    shape = tuple(a.shape)
    A = a.permute(2, 1, 0).contiguous()
    B = b.permute(2, 1, 0).contiguous()
    ctx = _lib.context(0)
    st = torch.cuda.Stream()
    out = {"workload": "3D FP32 512x512x256 particle volumes, shear + drift, interSize 32^3, searchMargin 8, "
                       "step 16 (C3)"}
    flop = flop_fft_per_ia(48, ndims=3)
    for alg in ("fft", "zncc"):
        p = setPIVParameters(corr_alg=alg, interSize=32, searchMargin=8, step=16)
        p.ndims = 3
        vf = get_vectorfield_size(shape, p)
        n = int(np.prod(vf))
        VF = torch.zeros(3 * n, dtype=torch.float64, device="cuda")

        def run():
            piv_device(ctx, p, shape, _lib.MQPIV_F32, A.data_ptr(), B.data_ptr(), None, _lib.MQPIV_F64,
                       VF.data_ptr(), None, st.cuda_stream, mode=mode)

        run()
        torch.cuda.synchronize()
        ctx.profile_begin()
        t0 = time.perf_counter()
        for _ in range(steps):
            run()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / steps
        kms = float(np.mean(ctx.profile_end()))
        out[alg] = {"n_ia": n, "ms_per_call": round(wall * 1e3, 3), "ia_per_s": round(n / wall, 1),
                    "kernel_ms": round(kms, 3), "achieved_tflops": round(n * flop / (kms * 1e-3) / 1e12, 3),
                    "flop_per_ia": round(flop)}
    if cpu:
        from oracle import oracle as O

        ha, hb = a.cpu().numpy(), b.cpu().numpy()
        n_s = 40 * nthreads
        lo = 7200 - n_s // 2  # IAs from the middle of the volume
        op = O.make_params("fft", 3, 32, 8, 16, False, np.float32)
        ts = []
        for r in range(4):  # one warm-up, then the median of three
            t0 = time.perf_counter()
            O.piv(ha, hb, op, mode=O.FFT, nthreads=nthreads, ia_begin=lo, ia_end=lo + n_s)
            if r:
                ts.append(time.perf_counter() - t0)
        dt = float(np.median(ts))
        out["cpu_baseline"] = {"value": round(n_s / dt, 1), "unit": "IA/s", "cores": nthreads, "kind": "port",
                               "sample": f"FFTCC, {n_s} IAs from the middle of the workload, Float64 oracle FFT "
                                         f"mode (48^3 transforms), OpenMP over IAs; median of 3 runs after a "
                                         f"warm-up ({dt:.2f} s each)"}
    return out


def batched_bench(mode: str, steps: int, seed: int = 4000, n_pairs: int = 16, side_range=(512, 4096),
                  sizes=(16, 32, 64)) -> dict:
    """Batched uneven masked NSQECC (one graph launch): n_pairs FP32 particle
    pairs with sides drawn from the multiples of 256 (of 64 below 512) in
    `side_range`, uniform shifts U[-6, 6]^2, a 20-ellipse mask (about 30% excluded)
    per pair, interSize drawn from `sizes` with searchMargin = step =
    interSize / 2.  The whole set runs as one captured HIP graph
    (integral-array-free masked measure: one grouped kernel launch per
    precision / transform length); the plain path runs the same pairs one call
    after another."""
    import torch

    from multi_quickpiv_amd import _lib, setPIVParameters, synth
    from multi_quickpiv_amd.piv import BatchGraph, DeviceRequest, piv_batch

    rng = np.random.default_rng(seed)
    reqs = []
    sides = []
    grid = 256 if side_range[0] >= 512 else 64
    for i in range(n_pairs):
        S = int(rng.choice(np.arange(side_range[0], side_range[1] + 1, grid)))
        isz = int(rng.choice(list(sizes)))
        d = tuple(rng.uniform(-6, 6, size=2))
        a, b = synth.particle_pair((S, S), synth.n_particles((S, S), 0.05), 1.2, synth.uniform(d), seed + i,
                                   device="cuda", dtype=torch.float32)
        m = synth.ellipse_mask((S, S), n=20, cover=0.3, seed=seed + 100 + i, device="cuda")
        p = setPIVParameters(corr_alg="nsqecc", interSize=isz, searchMargin=isz // 2, step=isz // 2)
        reqs.append(DeviceRequest(a, b, p, mask=m, mode=mode))
        sides.append((S, isz))
        del a, b, m
    n_ia = sum(int(np.prod(r.vf_size)) for r in reqs)
    ctx = _lib.context(0)
    st = torch.cuda.Stream()
    g = BatchGraph(reqs, ctx)
    for _ in range(2):
        g.launch(st.cuda_stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.launch(st.cuda_stream)
    torch.cuda.synchronize()
    t_graph = (time.perf_counter() - t0) / steps
    # plain per-pair path (one batched call per pair = the plain single-pair launch)
    for r in reqs:
        piv_batch([r], ctx, st.cuda_stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for r in reqs:
            piv_batch([r], ctx, st.cuda_stream)
    torch.cuda.synchronize()
    t_plain = (time.perf_counter() - t0) / steps
    g.close()
    return {"workload": f"{n_pairs} masked NSQECC FP32 pairs, sides {side_range[0]}-{side_range[1]} (multiples of {grid}), "
                        f"interSize {{{','.join(map(str, sizes))}}}, searchMargin = step = interSize/2, 20-ellipse masks"
                        + (" (C4)" if (n_pairs, tuple(side_range), tuple(sizes)) == (16, (512, 4096), (16, 32, 64)) else ""),
            "pairs": [{"side": S, "interSize": i} for S, i in sides], "ia_total": n_ia,
            "graph": {"ms_per_set": round(t_graph * 1e3, 3), "ia_per_s": round(n_ia / t_graph, 1)},
            "plain": {"ms_per_set": round(t_plain * 1e3, 3), "ia_per_s": round(n_ia / t_plain, 1)},
            "speedup": round(t_plain / t_graph, 3)}


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _mfma_executed_flop_per_ia(length):
    """MFMA flops one IA executes in the split-precision form (the hi/lo terms
    included), as rocprof's SQ_INSTS_VALU_MFMA_MOPS_F16 counted them."""
    f = ROOT / "profiles" / "mfma_split_2048_nsqecc_zncc.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith(f"mfma_flop_per_ia[{length}]"):
            return float(line.split("=")[1].split()[0])
    return None


def cpu_baseline_2d(ha, hb, jobs, nthreads: int, reps: int = 5) -> dict:
    """The restated reference CPU path (the Float64 oracle in FFT mode: the
    reference's per-IA structure, transforms at cs = good_size_real(ss)) on a
    bounded sample of the headline step: the same contiguous run of IAs from
    the middle of each of the four runs.  Timed single-threaded (the reference's
    serial IA loop, multi_quickPIV.jl:63) and OpenMP over IAs on the host's
    cores, each as the median of `reps` runs after one warm-up."""
    from oracle import oracle as O

    def timed(per_run: int, threads: int):
        ts = []
        for r in range(reps + 1):
            t = 0.0
            n = 0
            for j in jobs:
                op = O.make_params(j["alg"], 2, j["isz"], SM, STEP, True, np.float32)
                lo = j["n"] // 2 - per_run // 2
                t1 = time.perf_counter()
                O.piv(ha, hb, op, mode=O.FFT, nthreads=threads, ia_begin=lo, ia_end=lo + per_run)
                t += time.perf_counter() - t1
                n += per_run
            if r:  # the first run warms up
                ts.append(t)
        t_med = float(np.median(ts))
        return n / t_med, t_med, n

    one, t1, n1 = timed(200, 1)
    allc, ta, na = timed(200 * nthreads, nthreads)
    try:
        nproc = int(subprocess.run(["nproc"], capture_output=True, text=True).stdout.strip())
    except (OSError, ValueError):
        nproc = os.cpu_count() or 0
    return {"value": round(allc, 1), "unit": "IA/s", "cores": nthreads, "kind": "port",
            "sample": f"{na} IAs ({na // 4} from the middle of each of the four runs of the step), "
                      f"Float64 oracle FFT mode (per-IA transforms at the reference size), OpenMP over IAs on "
                      f"{nthreads} threads; median of {reps} runs after a warm-up ({ta:.2f} s each)",
            "single_thread": {"value": round(one, 1), "unit": "IA/s", "cores": 1,
                              "sample": f"{n1} IAs ({n1 // 4} per run), median of {reps} runs after a warm-up "
                                        f"({t1:.2f} s each)"},
            "nproc": nproc, "cpu_model": _cpu_model()}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--mode", default="auto", choices=["auto", "fp32", "fp64", "mfma"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one 256 MiB device-to-device copy (scale reference for the "
                         "FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--serve", action="store_true",
                    help="run only the serving benchmark (Poisson stream of mixed requests)")
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg of the default run")
    ap.add_argument("--no-3d", action="store_true", help="skip the 3-D leg of the default run")
    ap.add_argument("--only-3d", action="store_true", help="run only the 3-D leg (no CPU sample)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched masked leg of the default run")
    ap.add_argument("--only-batched", action="store_true", help="run only the batched masked leg")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-batch", type=int, default=32)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        return 2

    import torch

    from multi_quickpiv_amd import _lib, setPIVParameters, synth
    from multi_quickpiv_amd.piv import DeviceFrame, DeviceRequest, piv_batch, piv_device

    if args.only_3d:
        torch.cuda.set_device(0)
        if args.calib_copy:
            src = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")  # 256 MiB
            torch.empty_like(src).copy_(src)
            torch.cuda.synchronize()
            del src
        print(json.dumps({"volume_3d": volume_bench(args.mode, args.steps, False, 1)}))
        return 0

    if args.only_batched:
        torch.cuda.set_device(0)
        print(json.dumps({"batched_masked": batched_bench(args.mode, args.steps)}))
        return 0

    if args.serve:
        torch.cuda.set_device(0)
        res = serve_bench(args.serve_requests, [args.serve_rate, 0.0], 11, args.serve_slots, args.serve_batch,
                          mode=args.mode, log=lambda m: print(m, file=sys.stderr, flush=True))
        print(json.dumps({"metric": "served frame pairs/s with p50/p99 latency",
                          "value": res[f"rate_{int(args.serve_rate)}"]["served"]["pairs_per_s"],
                          "unit": "pairs/s", "n_gpus": 1, "higher_is_better": True, "data": "synthetic",
                          "config": {"workload": "serving"}, "serving": res}))
        return 0

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    a, b = synth.config_c2(SIZE, device="cuda")
    # the frame pair, column-major on the device once; the four runs are
    # requests on it, so one batch builds the second image's integral arrays
    # once for all four
    frame = DeviceFrame(a, b)
    ctx = _lib.context(0)
    # a dedicated (default-priority) stream: the kernels and the timing events
    # share it
    tstream = torch.cuda.Stream()
    torch.cuda.set_stream(tstream)
    stream = tstream.cuda_stream

    jobs = []
    for alg, isz in RUNS:
        p = setPIVParameters(corr_alg=alg, interSize=isz, searchMargin=SM, step=STEP, computeSN=True)
        r = DeviceRequest.from_frame(frame, p, mode=args.mode)
        n = int(np.prod(r.vf_size))
        jobs.append(dict(alg=alg, isz=isz, p=p, vf=r.vf_size, n=n, req=r, VF=r.VF, SN=r.SN,
                         form=form_of(args.mode, alg, isz, True)))
    reqs = [j["req"] for j in jobs]

    def step():
        piv_batch(reqs, ctx, stream)

    def launch(j):  # one run as its own call (the per-call comparison)
        piv_device(ctx, j["p"], (SIZE, SIZE), _lib.MQPIV_F32, frame.a.data_ptr(), frame.b.data_ptr(), None,
                   _lib.MQPIV_F64, j["VF"].data_ptr(), j["SN"].data_ptr(), stream, mode=args.mode)

    for _ in range(args.warmup):
        step()
        for j in jobs:
            launch(j)
    torch.cuda.synchronize()
    if args.calib_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)  # 256 MiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    step_ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(args.steps)]
    torch.cuda.synchronize()
    ctx.profile_begin()
    t0 = time.perf_counter()
    for s in range(args.steps):
        step_ev[s][0].record()
        step()
        step_ev[s][1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kt = ctx.profile_end()  # correlation-kernel durations, launch order

    n_step = sum(j["n"] for j in jobs)
    value = n_step * args.steps / wall
    assert len(kt) == args.steps * len(jobs), len(kt)
    kern_ms = [float(np.mean(kt[k::len(jobs)])) for k in range(len(jobs))]
    step_ms = np.array([step_ev[s][0].elapsed_time(step_ev[s][1]) for s in range(args.steps)])

    # the same four runs as four separate calls (each builds its own integral
    # arrays), for comparison
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for s in range(args.steps):
        for j in jobs:
            launch(j)
    torch.cuda.synchronize()
    wall_calls = time.perf_counter() - t1

    # the interSize 32 runs in the two GPU transform forms, one call each: FP32
    # radix transforms in registers and LDS, and the split-precision
    # matrix-core form (the Float64 CPU oracle is the cpu_baseline leg)
    forms_is32 = {"what": "the four runs, one call each, correlation-kernel time (events)"}
    for j in jobs:
        res = {}
        for m in ("fp32", "mfma"):
            def call(m=m, j=j):
                piv_device(ctx, j["p"], (SIZE, SIZE), _lib.MQPIV_F32, frame.a.data_ptr(), frame.b.data_ptr(),
                           None, _lib.MQPIV_F64, j["VF"].data_ptr(), j["SN"].data_ptr(), stream, mode=m)
            for _ in range(3):
                call()
            torch.cuda.synchronize()
            nc = max(5, min(args.steps, 30))
            ctx.profile_begin()
            t2 = time.perf_counter()
            for _ in range(nc):
                call()
            torch.cuda.synchronize()
            w2 = time.perf_counter() - t2
            kms = float(np.mean(ctx.profile_end()))
            res[m] = {"kernel_ms": round(kms, 4), "ia_per_s": round(j["n"] / (kms * 1e-3), 1),
                      "ia_per_s_per_call": round(j["n"] * nc / w2, 1)}
        cs32 = j["isz"] + 2 * SM
        ach = flop_dft_per_ia(cs32) * j["n"] / (res["mfma"]["kernel_ms"] * 1e-3) / 1e12
        res["mfma"]["roofline"] = {"bound": "mfma", "achieved": round(ach, 3), "peak": PEAK_FP16_MFMA_TFLOPS,
                                   "unit": "TFLOP/s", "frac": round(ach / PEAK_FP16_MFMA_TFLOPS, 5),
                                   "flop_per_ia": round(flop_dft_per_ia(cs32))}
        ex = _mfma_executed_flop_per_ia(cs32)
        if ex is not None:
            exr = ex * j["n"] / (res["mfma"]["kernel_ms"] * 1e-3) / 1e12
            res["mfma"]["executed"] = {"flop_per_ia": ex, "tflops": round(exr, 2),
                                       "frac": round(exr / PEAK_FP16_MFMA_TFLOPS, 4)}
        res["mfma_over_fp32"] = round(res["fp32"]["kernel_ms"] / res["mfma"]["kernel_ms"], 3)
        forms_is32[f'{j["alg"]}{j["isz"]}'] = res

    # the batched split-precision form against the plain per-frame FP32 path
    # at interSize 32: the step's IA-32 runs (one batched call, kernel time)
    # and the same runs as plain FP32 calls (one call each, wall time per call)
    i32 = [k for k, j in enumerate(jobs) if j["isz"] == 32]
    n32 = sum(jobs[k]["n"] for k in i32)
    rate_b = n32 / (sum(kern_ms[k] for k in i32) * 1e-3)
    rate_p = n32 / sum(jobs[k]["n"] / forms_is32[f'{jobs[k]["alg"]}{jobs[k]["isz"]}']["fp32"]["ia_per_s_per_call"]
                       for k in i32)
    split_vs_plain = {"what": "IA/s of the batched split-precision form (the step's interSize-32 runs, ZNCC+SN "
                          "with its FP64 recomputation and NSQECC) over the plain per-frame FP32 path (mode "
                          "fp32, one call per run) on the 2048^2 pair",
                  "batched_split_ia_per_s": round(rate_b, 1), "plain_fp32_ia_per_s": round(rate_p, 1),
                  "ratio": round(rate_b / rate_p, 3), "target": 5.0}

    # roofline of the dominant run (its launch: for "split-f16+f64" the
    # split-precision kernel, then the FP64 kernel over the IAs it flagged).
    # The split-precision form is bound by the matrix cores: its algorithmic
    # flops are the dense-DFT products' (flop_dft_per_ia), against the dense
    # FP16 peak; the radix forms by the vector unit (flop_fft_per_ia)
    kd = int(np.argmax(kern_ms))
    jd = jobs[kd]
    cs = jd["isz"] + 2 * SM  # reference transform size (good_size_real is the identity here)
    split = jd["form"].startswith("split-f16")
    flop_ia = flop_dft_per_ia(cs) if split else flop_fft_per_ia(cs)
    achieved = flop_ia * jd["n"] / (kern_ms[kd] * 1e-3) / 1e12
    peak = PEAK_FP16_MFMA_TFLOPS if split else (PEAK_FP64_TFLOPS if jd["form"] == "f64" else PEAK_FP32_TFLOPS)
    traffic = None
    tfile = ROOT / "profiles" / "piv2d_2048_zncc_nsqecc_is32_is64.md"
    if tfile.exists():
        for line in tfile.read_text().splitlines():
            if line.startswith(f"traffic_bytes[{jd['alg']}{jd['isz']}_{jd['form']}]"):
                traffic = float(line.split("=")[1].split()[0])

    # parity spot check of the timed outputs against the resident data (the
    # tests hold the full parity; here: finite vectors, the vortex is resolved)
    VF0 = jobs[1]["VF"].view(jobs[1]["vf"][1], jobs[1]["vf"][0], 2).cpu().numpy()
    finite = bool(np.isfinite(VF0).all())

    breakdown = []
    for k, j in enumerate(jobs):
        breakdown.append({"alg": j["alg"], "interSize": j["isz"], "n_ia": j["n"],
                          "transforms": j["form"],
                          "kernel_ms": round(kern_ms[k], 4),
                          "ia_per_s": round(j["n"] / (kern_ms[k] * 1e-3), 1)})

    nthreads = args.cpu_threads or min(16, os.cpu_count() or 1)
    volume = None
    if not args.no_3d:
        volume = volume_bench(args.mode, 3, not args.no_cpu, nthreads)

    batched = batched_small = None
    if not args.no_batched:
        batched = batched_bench(args.mode, 3)
        # where batching pays: many small pairs, whose launches alone leave most CUs idle
        batched_small = batched_bench(args.mode, 10, seed=4500, n_pairs=64, side_range=(128, 512), sizes=(16, 32))

    serving = None
    if not args.no_serve:
        serving = serve_bench(args.serve_requests, [args.serve_rate, 0.0], 11, args.serve_slots,
                              args.serve_batch, mode=args.mode)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_2d(a.cpu().numpy(), b.cpu().numpy(), jobs, nthreads)

    # the arithmetic of the step: the forms its runs use ("split-f16": hi + lo FP16
    # products accumulated in FP32; every form keeps FP64 window sums and
    # denominators)
    dtype = "+".join(sorted({j["form"] for j in jobs}))
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "IA/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {
            "workload": "2D ZNCC+NSQECC, FP32 2048x2048 Lamb-Oseen particle pair, interSize 32 and 64, "
                        "searchMargin 16, step 16, computeSN; the four runs as one batched call per pair",
            "ia_per_step": n_step,
            "precision_mode": args.mode,
        },
        "latency_ms": {"p50": round(float(np.percentile(step_ms, 50)), 4),
                       "p99": round(float(np.percentile(step_ms, 99)), 4),
                       "what": "one frame pair through all four runs"},
        "roofline": {"bound": "mfma" if split else "valu",
                     "kernel": (f"piv2d_wave_kernel {jd['alg']} interSize {jd['isz']}" if split else
                                f"piv2d_kernel {jd['alg']} interSize {jd['isz']}") + f" ({jd['form']} transforms)",
                     "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4), "traffic": traffic,
                     "flop_per_ia": round(flop_ia), "ia_per_launch": jd["n"],
                     "kernel_ms": round(kern_ms[kd], 4)},
        "breakdown": breakdown,
        "per_call": {"value": round(n_step * args.steps / wall_calls, 1),
                     "ms_per_step": round(wall_calls * 1e3 / args.steps, 4),
                     "what": "the same four runs as four separate PIV calls per step"},
        "transform_forms": forms_is32,
        "split_vs_plain_fp32": split_vs_plain,
        "outputs_finite": finite,
        "volume_3d": volume,
        "batched_masked": batched,
        "batched_masked_small": batched_small,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
