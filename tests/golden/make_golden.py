"""Writes tests/golden/golden.json: per-picture hashes of the reference decoder's
own reconstruction on generated syntax.

The reference's reconstruction sources (jfu222/h264_video_decoder_demo, compiled
unmodified by oracle/Makefile into oracle/_ref/) are driven by oracle/ref_harness.cpp
on the syntax of csrc/gen.cpp.  For each case the fixture records the generator's
digest (so a changed generator is caught) and SHA-256 of every output picture's Y,
Cb and Cr planes, before and after deblocking, and of the BGR24 image the
reference's convertYuv420pToBgr24 makes of it.  The reference has no tests or
golden vectors of its own for this path; these are the pins.

Run in a container that has /root/reference:  python tests/golden/make_golden.py [case ...]
(with case names, only those cases are recomputed and the others are kept).
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")

# name -> (generator config, overrides as generator parameter names)
CASES = {
    # 352x288, I + P, the plain-path workload
    "cif_ip": (1, {}),
    "cif_ip_seed2": (1, {"seed": 2}),
    "cif_ip_seed3": (1, {"seed": 3}),
    # config-1 syntax over more frames with I_PCM, chroma QP offsets, idc 2
    "cif_pcm_cqp": (1, {"width_mbs": 30, "height_mbs": 20, "n_frames": 6, "ipcm_permille": 30,
                        "chroma_qp_index_offset": -4, "second_chroma_qp_index_offset": 3,
                        "disable_deblocking_filter_idc": 2}),
    # High-profile IPBBPBBP syntax at a reduced size
    "high_small": (2, {"width_mbs": 40, "height_mbs": 24}),
    "high_small_seed7": (2, {"width_mbs": 23, "height_mbs": 17, "seed": 7}),
    # implicit bi-prediction weights, explicit P weights
    "high_implicit": (2, {"width_mbs": 24, "height_mbs": 16, "weighted_pred_flag": 1, "weighted_bipred_idc": 2,
                          "chroma_qp_index_offset": 2, "second_chroma_qp_index_offset": -3}),
    # default weights everywhere, deblocking off, I_PCM
    "high_default_nodbk": (2, {"width_mbs": 24, "height_mbs": 16, "weighted_pred_flag": 0, "weighted_bipred_idc": 0,
                               "disable_deblocking_filter_idc": 1, "ipcm_permille": 20}),
    # one reference, no B pictures, single-MB-wide and single-MB-tall pictures
    "narrow": (2, {"width_mbs": 1, "height_mbs": 12, "max_refs": 1, "b_frames": 0}),
    "flat": (2, {"width_mbs": 14, "height_mbs": 1, "max_refs": 2}),
    # the full 1920x1088 High-profile workload (8 pictures)
    "high_1080p": (2, {}),
    # the other full picture sizes of the multi-stream and serving workloads, config-2
    # syntax, I P B B
    "high_720p": (2, {"width_mbs": 80, "height_mbs": 45, "n_frames": 4, "seed": 0x5EED0004 + 16}),
    "high_2160p": (2, {"width_mbs": 240, "height_mbs": 135, "n_frames": 4, "seed": 0x5EED0004 + 56}),
    # MBAFF frames (config 3 syntax: field / frame MB pairs 50/50, 4x4 transform): reduced
    # size I P B B P B B P; 8x8 transforms (the 8x8 field scan) with I_PCM and idc 2; a
    # two-pair-wide picture (edges on both sides of every pair); the full 1920x1088 stream
    # (each frame buffer at least 256 KiB, which glibc always maps on pages of its own:
    # ref_harness checks that the bytes the reference reads past it are zero)
    "mbaff_small": (3, {"width_mbs": 32, "height_mbs": 22, "seed": 2}),
    "mbaff_t8_pcm": (3, {"width_mbs": 28, "height_mbs": 26, "n_frames": 6, "seed": 4, "t8x8": 1,
                         "ipcm_permille": 30, "disable_deblocking_filter_idc": 2}),
    "mbaff_tall": (3, {"width_mbs": 2, "height_mbs": 344, "n_frames": 4, "seed": 5}),
    "mbaff_1080p": (3, {}),
    # constrained intra prediction: intra MBs predict from intra neighbours only, and an
    # inter neighbour makes the intra-mode prediction DC
    "high_small_cip": (2, {"width_mbs": 40, "height_mbs": 24, "constrained_intra_pred": 1, "seed": 12}),
    "high_odd_cip": (2, {"width_mbs": 23, "height_mbs": 17, "constrained_intra_pred": 1, "seed": 13}),
    # spatial direct prediction in the B pictures: B_Skip, B_Direct_16x16 and B_8x8 with
    # B_Direct_8x8 sub-macroblocks beside every other kind of MB
    "high_direct": (2, {"width_mbs": 40, "height_mbs": 24, "direct_permille": 400, "seed": 21}),
    "high_direct_odd": (2, {"width_mbs": 23, "height_mbs": 17, "direct_permille": 600, "seed": 22}),
    # I_PCM co-located MBs: the reference reads their zeroed motion as index 0, vector 0
    "high_direct_pcm": (2, {"width_mbs": 30, "height_mbs": 20, "direct_permille": 500, "ipcm_permille": 150,
                            "seed": 23}),
}

_CLI = {"n_frames": "--frames", "width_mbs": "--width", "height_mbs": "--height", "seed": "--seed",
        "ipcm_permille": "--ipcm", "disable_deblocking_filter_idc": "--deblock", "max_refs": "--refs",
        "b_frames": "--bframes", "t8x8": "--t8x8", "mbaff": "--mbaff", "constrained_intra_pred": "--cip",
        "direct_permille": "--direct"}


def harness_args(config: int, ov: dict) -> list:
    args = ["--config", str(config)]
    for k, v in ov.items():
        if k in _CLI:
            args += [_CLI[k], str(v)]
    if "weighted_pred_flag" in ov or "weighted_bipred_idc" in ov:
        args += ["--wp", f"{ov.get('weighted_pred_flag', 0)},{ov.get('weighted_bipred_idc', 1)}"]
    if "chroma_qp_index_offset" in ov or "second_chroma_qp_index_offset" in ov:
        args += ["--cqp", f"{ov.get('chroma_qp_index_offset', 0)},{ov.get('second_chroma_qp_index_offset', 0)}"]
    return args


def plane_hashes(buf: bytes, W: int, H: int):
    ny, nc = W * H * 256, W * H * 64
    return [hashlib.sha256(buf[:ny]).hexdigest(), hashlib.sha256(buf[ny:ny + nc]).hexdigest(),
            hashlib.sha256(buf[ny + nc:]).hexdigest()]


def run_case(config: int, ov: dict):
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([HARNESS, *harness_args(config, ov), "--predeblock", "--bgr", "--out", d],
                           capture_output=True, text=True, check=True)
        info = json.loads(r.stdout.strip().splitlines()[-1])
        n = info["pictures"]
        files = sorted(os.listdir(d))
        W, H = None, None
        out = {"digest": info["digest"], "pictures": n, "pic": [], "pre": [], "bgr": []}
        for i in range(n):
            for kind in ("pre", "pic"):
                buf = open(os.path.join(d, f"{kind}_{i:03d}.yuv"), "rb").read()
                if W is None:
                    # sizes come from the generator defaults + overrides
                    from h264_video_decoder_demo_amd import Generator
                    g = Generator(config, **ov)
                    W, H = g.width_mbs, g.height_mbs
                assert len(buf) == W * H * 384, (kind, i, len(buf))
                out[kind].append(plane_hashes(buf, W, H))
            bgr = open(os.path.join(d, f"bgr_{i:03d}.raw"), "rb").read()
            assert len(bgr) == (W * 16 * 3 + 3) // 4 * 4 * H * 16
            out["bgr"].append(hashlib.sha256(bgr).hexdigest())
        out["width_mbs"], out["height_mbs"] = W, H
        assert info["mv_mismatches"] == 0, info
        del files
        return out


def main():
    if not os.path.exists(HARNESS):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref", "-j8"], check=True)
    sys.path.insert(0, ROOT)
    path = os.path.join(HERE, "golden.json")
    only = sys.argv[1:]
    golden = {}
    if only:
        with open(path) as f:
            golden = json.load(f)
    for name, (config, ov) in CASES.items():
        if only and name not in only:
            continue
        golden[name] = {"config": config, "overrides": ov, **run_case(config, ov)}
        print(name, golden[name]["digest"], golden[name]["pictures"], flush=True)
    with open(path, "w") as f:
        json.dump(golden, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
