#!/usr/bin/env python
"""How much the reference's outputs depend on the values of its aliased
bispectrum lookups, and how far the oracle is from the one reference-made
output there is (test infrastructure: a measurement with the oracle, not part
of the product).

get_element forms an address with calc_offset's size_t arithmetic even when the
canonical (u, v) of a T3/T9 lookup lies outside the stored v range
(include/bispectrum.h:482-512,324-349); no reference output can be produced
here to pin the values read there.  This script reruns the oracle's phase
recovery with those lookups left out of each mean (alias_mode 1, a probe, not
the reference) and reports how far flags, phases and image move:

  * binary128: the 40-frame 128x128 uint8 binary-star stack, b = 32, p = 64;
  * triple256: the first `--c2-frames` frames of the 256x256 triple-star stack,
    b = 48, p = 128;
  * hu940ani: the reference's README example (-b 32 -p 64 -c r), with the
    distance of each variant from the reference's own reco_image.png, in 16-bit
    levels (LSB) after the CLI's mapping (pixel (j, i) shows index
    (j - X/2, i - Y/2), value * 65535 truncated).

    python oracle/alias_sensitivity.py [--c2-frames 16]
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(REPO, "speckle-stacker_amd", "python"))

import oracle as O  # noqa: E402


def recover(frames: np.ndarray, b: int, p: int, bispec_f32: bool, alias_mode: int):
    n, Y, X = frames.shape
    B, P = O.power_and_bispectrum(frames.astype(np.float64), b, bispec_f32=bispec_f32)
    r = O.reconstruct_phases(O.desc(X, Y, b), B, X, Y, p, alias_mode=alias_mode)
    img = O.synthesize(P, r.phases, p)
    return r, img


def compare(a, b) -> dict:
    (ra, ia), (rb, ib) = a, b
    both = ra.flags.astype(bool) & rb.flags.astype(bool)
    return {"flags_changed": int(np.sum(ra.flags != rb.flags)), "flagged": int(ra.flags.sum()),
            "phase_max_abs_diff": float(np.abs(ra.phases[both] - rb.phases[both]).max()),
            "phase_rms_diff": float(np.sqrt(np.mean(np.abs(ra.phases[both] - rb.phases[both]) ** 2))),
            "image_rel_l2": float(np.linalg.norm(np.abs(ia) - np.abs(ib)) / np.linalg.norm(np.abs(ib))),
            "lookups": int(rb.stats["lookups"]), "aliased_lookups": int(rb.stats["aliased"])}


def centred(a):
    return np.roll(np.roll(a, a.shape[0] // 2, axis=0), a.shape[1] // 2, axis=1)


def grey16(x):
    return np.trunc(np.asarray(x, dtype=np.float64) * 65535.0).astype(np.int64) & 0xFFFF


def lsb_stats(img: np.ndarray, ref: np.ndarray) -> dict:
    ours = grey16(centred(np.abs(img)))
    d = np.abs(ours - ref)
    return {"corr": float(np.corrcoef(ours.ravel(), ref.ravel())[0, 1]), "median_lsb": float(np.median(d)),
            "p99_lsb": float(np.percentile(d, 99)), "max_lsb": int(d.max()),
            "within_1_lsb": float(np.mean(d <= 1)), "peak": [int(v) for v in np.unravel_index(ours.argmax(), ours.shape)]}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--c2-frames", type=int, default=16)
    args = ap.parse_args()
    import smip_amd as S
    from PIL import Image, ImageSequence

    out = {}
    for name, n in (("binary128", 40), ("triple256", args.c2_frames)):
        W = S.WORKLOADS[name]
        fr = S.synth_frames(S.synth_params(name), 0, n)
        base = recover(fr, W["b"], W["p"], False, 0)
        probe = recover(fr, W["b"], W["p"], False, 1)
        out[name] = {"frames": n, "b": W["b"], "p": W["p"], "aliased_left_out_vs_reference": compare(probe, base)}
        print(name, json.dumps(out[name]), flush=True)

    gif = os.path.join(REPO, "tests", "golden", "hu940ani.gif")
    png = os.path.join(REPO, "tests", "golden", "hu940ani_reco_image.png")
    im = Image.open(gif)
    fr = np.stack([np.asarray(f.convert("RGB")) for f in ImageSequence.Iterator(im)])[..., 0]
    ref = np.asarray(Image.open(png)).astype(np.int64)
    res = {}
    for label, f32, mode in (("fp64_bispectrum", False, 0), ("fp32_bispectrum_as_cli", True, 0),
                             ("fp32_bispectrum_aliased_left_out", True, 1)):
        res[label] = recover(fr, 32, 64, f32, mode)
    out["hu940ani"] = {"vs_reference_png": {k: lsb_stats(v[1], ref) for k, v in res.items()},
                       "aliased_left_out_vs_reference": compare(res["fp32_bispectrum_aliased_left_out"],
                                                                res["fp32_bispectrum_as_cli"]),
                       "fp32_vs_fp64_bispectrum": compare(res["fp32_bispectrum_as_cli"], res["fp64_bispectrum"])}
    print("hu940ani", json.dumps(out["hu940ani"]), flush=True)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
