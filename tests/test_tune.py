"""Host logic of the lanes-per-line autotuner (no GPU): candidate validity, the
choice among agreeing candidates, and the rewrite of the LPL header."""
from __future__ import annotations

import re

from multi_quickpiv_amd import tune


def _times(base: float, **over):
    t = {}
    for tag, lens in tune.LENGTHS.items():
        for L in lens:
            for alg in ("fft", "nsqecc", "zncc"):
                t[f"{tag}_{L}_{alg}"] = base
    t.update(over)
    return t


def test_valid_shapes():
    assert tune._valid("f32", 64, 4) and tune._valid("f32", 32, 1)
    assert not tune._valid("f32", 48, 1)  # 48 values per lane
    assert not tune._valid("f64", 32, 1)  # FP64 holds at most 24 per lane
    assert not tune._valid("f32", 16, 8)  # 2 per lane
    assert not tune._valid("f32", 48, 32)


def test_choose_fastest_agreeing_candidate():
    times = {"1": _times(1.0), "2": _times(2.0), "4": _times(3.0), "8": _times(4.0)}
    times["4"]["f32_64_fft"] = times["4"]["f32_64_nsqecc"] = times["4"]["f32_64_zncc"] = 0.1
    times["2"]["f32_32_zncc"] = 0.0  # fastest, but its field disagrees
    times["1"]["f32_32_fft"] = 9.0
    agree = {k: ["1", "2", "4", "8"] for k in times["1"]}
    agree["f32_32_zncc"] = ["1", "4", "8"]
    best = tune.choose({"times_ms": times, "agree": agree})
    assert best[("f32", 64)] == 4
    assert best[("f32", 32)] == 4  # 1 is slower in sum, 2 disagrees
    assert best[("f64", 96)] == 4  # 24 values a lane is the most FP64 allows
    assert best[("f32", 16)] == 1


def test_header_has_every_entry():
    text = tune.LPL_HEADER.read_text()
    for tag, lens in tune.LENGTHS.items():
        for L in lens:
            m = re.search(rf"#define MQPIV_LPL_{tag.upper()}_{L} (\d+)", text)
            assert m and tune._valid(tag, L, int(m.group(1)))
