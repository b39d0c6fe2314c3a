"""Autotuner for the fused 2-D kernel's shape: lanes per transform line (LPL)
for every (scalar type, transform length), measured on the GPU.

    python -m multi_quickpiv_amd.tune build            # CPU: one library per candidate LPL
    python -m multi_quickpiv_amd.tune measure [out]    # GPU: time every instance in every library
    python -m multi_quickpiv_amd.tune apply [out]      # write the fastest LPL per (type, L) into
                                                       # csrc/mqpiv_lpl2d.h (then rebuild)

A candidate library compiles every (type, L) with the same LPL (lengths where
that LPL is not valid keep their current entry), so four builds cover the
whole table.  `measure` runs each library in its own process (MQPIV_LIB) on a
2048 x 2048 FP32 particle pair with interSize L/2, searchMargin L/4 (search
size = L) and step L/8, for FFTCC, NSQECC and ZNCC with SN, and records the median
kernel time of 25 launches from the library's own event timing, and the
vector field.  `measure` then drops, per instance, a candidate whose vector field does
not agree with the other candidates' (more than 0.1 % of vectors off by more
than 1e-3 px: the lane split changes only the transforms' rounding); `apply`
picks per (type, L) the LPL with the smallest summed kernel time over the algorithms.
"""
from __future__ import annotations

import json
import os
import re
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
TUNE_DIR = PKG / "_tune"
LPL_HEADER = PKG / "csrc" / "mqpiv_lpl2d.h"
LENGTHS = {"f32": [16, 32, 48, 64, 96, 128], "f64": [16, 32, 48, 64, 96]}
CANDIDATES = [1, 2, 4, 8]
REPS = 25


def _valid(tag: str, L: int, lpl: int) -> bool:
    """Shapes the kernel supports and that fit in registers: a line of L
    samples over LPL lanes with at most 32 complex values per lane (24 for
    FP64), and at most 1024 lanes per IA."""
    if L % lpl or L * lpl > 1024:
        return False
    per_lane = L // lpl
    return per_lane <= (32 if tag == "f32" else 24) and per_lane >= 4


def build(jobs: int = 8) -> list[Path]:
    from .build import build_lib

    TUNE_DIR.mkdir(exist_ok=True)
    out = []
    for lpl in CANDIDATES:
        flags = []
        for tag, lens in LENGTHS.items():
            for L in lens:
                if _valid(tag, L, lpl):
                    flags.append(f"-DMQPIV_LPL_{tag.upper()}_{L}={lpl}")
        lib = TUNE_DIR / f"lib_lpl{lpl}.so"
        build_lib(jobs=jobs, out=lib, extra=flags)
        (TUNE_DIR / f"lib_lpl{lpl}.json").write_text(json.dumps(flags))
        out.append(lib)
        print(f"built {lib}", flush=True)
    return out


def _measure_one(lpl: int) -> dict:
    """Runs in a child process with MQPIV_LIB set: kernel times for every
    (type, L, algorithm) of that library; the vector fields go to an .npz."""
    import numpy as np
    import torch

    from . import _lib, setPIVParameters, synth
    from .params import get_vectorfield_size
    from .piv import piv_device

    torch.cuda.set_device(0)
    a, b = synth.config_c2(2048, device="cuda")
    A, B = a.t().contiguous(), b.t().contiguous()
    ctx = _lib.context(0)
    st = torch.cuda.Stream()
    res, fields = {}, {}
    for tag, lens in LENGTHS.items():
        mode = "fp32" if tag == "f32" else "fp64"
        for L in lens:
            for alg in ("fft", "nsqecc", "zncc"):
                isz, sm = L // 2, L // 4
                p = setPIVParameters(corr_alg=alg, interSize=isz, searchMargin=sm, step=max(4, L // 8),
                                     computeSN=True)
                p.ndims = 2
                vf = get_vectorfield_size((2048, 2048), p)
                n = int(np.prod(vf))
                VF = torch.zeros(2 * n, dtype=torch.float64, device="cuda")
                SN = torch.zeros(n, dtype=torch.float64, device="cuda")

                def run():
                    piv_device(ctx, p, (2048, 2048), _lib.MQPIV_F32, A.data_ptr(), B.data_ptr(), None,
                               _lib.MQPIV_F64, VF.data_ptr(), SN.data_ptr(), st.cuda_stream, mode=mode)

                try:
                    for _ in range(3):
                        run()
                    torch.cuda.synchronize()
                    ctx.profile_begin()
                    for _ in range(REPS):
                        run()
                    torch.cuda.synchronize()
                    kt = ctx.profile_end()
                    res[f"{tag}_{L}_{alg}"] = float(np.median(kt)) if kt else None
                    fields[f"{tag}_{L}_{alg}"] = VF.cpu().numpy()
                except _lib.MqpivError as e:  # a shape the build does not support (e.g. LDS)
                    res[f"{tag}_{L}_{alg}"] = None
                    print(f"{tag} L={L} {alg}: {e}", file=sys.stderr)
    np.savez(TUNE_DIR / f"vf_lpl{lpl}.npz", **fields)
    return res


def measure(out: Path) -> dict:
    results = {}
    # two sweeps in opposite orders, the faster of the two per instance, so a
    # drift of the clocks during the run favours no candidate
    for lpl in CANDIDATES + CANDIDATES[::-1]:
        lib = TUNE_DIR / f"lib_lpl{lpl}.so"
        if not lib.exists():
            continue
        env = dict(os.environ, MQPIV_LIB=str(lib))
        r = subprocess.run([sys.executable, "-m", "multi_quickpiv_amd.tune", "_child", str(lpl)], env=env,
                           capture_output=True, text=True, timeout=900)
        if r.returncode != 0:
            print(r.stderr[-2000:], file=sys.stderr)
            raise SystemExit(f"candidate lpl {lpl} failed")
        got = json.loads(r.stdout.strip().splitlines()[-1])
        prev = results.setdefault(str(lpl), got)
        for k, v in got.items():
            if v is not None and (prev.get(k) is None or v < prev[k]):
                prev[k] = v
        print(f"lpl {lpl}: done", flush=True)
    agree = {k: sorted(v) for k, v in _agreeing(results).items()}
    for lpl in results:
        (TUNE_DIR / f"vf_lpl{lpl}.npz").unlink(missing_ok=True)
    doc = {"times_ms": results, "agree": agree}
    out.write_text(json.dumps(doc, indent=1))
    return doc


def _agreeing(results: dict) -> dict:
    """Per instance key, the candidates whose vector field agrees with at least
    half of the other candidates' (all of them when there is one)."""
    import numpy as np

    fields = {}
    for lpl in results:
        f = TUNE_DIR / f"vf_lpl{lpl}.npz"
        if f.exists():
            with np.load(f) as z:
                fields[lpl] = {k: z[k] for k in z.files}
    ok = {}
    keys = {k for r in results.values() for k in r}
    for key in keys:
        have = [l for l in results if key in fields.get(l, {})]
        good = set()
        for l in have:
            agree = sum(
                float(np.mean(np.abs(fields[l][key] - fields[m][key]) > 1e-3)) <= 1e-3 for m in have if m != l
            )
            if len(have) == 1 or agree >= (len(have) - 1) / 2:
                good.add(l)
        ok[key] = good
    return ok


def choose(doc: dict) -> dict:
    """Per (type, L) the LPL with the smallest summed kernel time among the
    candidates whose vector fields agree."""
    results, ok = doc["times_ms"], doc["agree"]
    best = {}
    for tag, lens in LENGTHS.items():
        for L in lens:
            scores = {}
            for lpl, r in results.items():
                keys = [f"{tag}_{L}_{alg}" for alg in ("fft", "nsqecc", "zncc")]
                if not _valid(tag, L, int(lpl)) or not all(lpl in ok.get(k, ()) for k in keys):
                    continue
                ts = [r.get(k) for k in keys]
                if all(t is not None for t in ts):
                    scores[int(lpl)] = sum(ts)
            if scores:
                best[(tag, L)] = min(scores, key=scores.get)
    return best


def apply(results_path: Path) -> dict:
    best = choose(json.loads(results_path.read_text()))
    text = LPL_HEADER.read_text()
    for (tag, L), lpl in best.items():
        text = re.sub(rf"#define MQPIV_LPL_{tag.upper()}_{L} \d+", f"#define MQPIV_LPL_{tag.upper()}_{L} {lpl}", text)
    LPL_HEADER.write_text(text)
    for (tag, L), lpl in sorted(best.items()):
        print(f"{tag} L={L}: LPL {lpl}")
    return best


def main(argv: list[str]) -> int:
    cmd = argv[0] if argv else ""
    out = Path(argv[1]) if len(argv) > 1 and cmd != "_child" else TUNE_DIR / "results.json"
    if cmd == "build":
        build()
    elif cmd == "measure":
        TUNE_DIR.mkdir(exist_ok=True)
        measure(out)
    elif cmd == "apply":
        apply(out)
    elif cmd == "_child":
        print(json.dumps(_measure_one(int(argv[1]))))
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    raise SystemExit(main(sys.argv[1:]))
