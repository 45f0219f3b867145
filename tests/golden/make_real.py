"""Writes the real-stream fixtures: prefixes of the reference's own sample streams and
the reference decoder's output for each.

  tests/golden/heavyhand_cavlc_15.h264   the first 15 pictures (whole NAL units) of
      data/HeavyHand_1080p.B_frames_4.no_cabac.no_tff.h264: High profile, CAVLC,
      1920x1088 frames, B pyramids, explicit (P) and implicit (B) weighted prediction,
      4 reference frames, reference list modification
  tests/golden/heavyhand_cabac_10.h264   the first 10 pictures of
      data/HeavyHand_1080p.B_frames.cabac.h264: the same content coded with CABAC
  tests/golden/gop121_12.h264            the first 12 pictures of data/gop121.naluCnt453.h264:
      Baseline profile, CAVLC, 9 slices per picture, disable_deblocking_filter_idc 2; the
      last slice of its first picture is damaged (its data fails at MB 7713), which the
      reference decoder survives by abandoning the slice there
  tests/golden/<name>.json               per output picture of the reference decoder
      (oracle/_ref/ref_decode: every reference translation unit but main.cpp, compiled
      unmodified by `make -C oracle decode`), in its output order: PicOrderCnt, slice
      type, decoding index, SHA-256 of the Y / Cb / Cr planes

Run in a container that has the reference:  python3 tests/golden/make_real.py
"""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
HERE = os.path.dirname(os.path.abspath(__file__))
DATA = "/root/reference/data"
STREAMS = [("HeavyHand_1080p.B_frames_4.no_cabac.no_tff.h264", 15, "heavyhand_cavlc_15"),
           ("HeavyHand_1080p.B_frames.cabac.h264", 10, "heavyhand_cabac_10"),
           ("gop121.naluCnt453.h264", 12, "gop121_12")]


def prefix(data: bytes, n_pictures: int) -> bytes:
    """The stream up to the end of the last slice NAL unit of its n-th picture (a picture
    starts with a slice whose first_mb_in_slice, the first ue(v) of the header, is 0: bit
    '1')."""
    starts, i = [], 0
    while True:
        j = data.find(b"\x00\x00\x01", i)
        if j < 0:
            break
        starts.append(j)
        i = j + 3
    pics, end = 0, None
    for k, s in enumerate(starts):
        if data[s + 3] & 31 in (1, 5):
            if data[s + 4] & 0x80:
                if pics == n_pictures:
                    return data[:end]
                pics += 1
            end = starts[k + 1] if k + 1 < len(starts) else len(data)
    if pics == n_pictures:
        return data[:end]
    raise ValueError("the stream has fewer pictures")


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "decode"], check=True, capture_output=True)
    for src, n_pictures, name in STREAMS:
        with open(os.path.join(DATA, src), "rb") as f:
            data = prefix(f.read(), n_pictures)
        path = os.path.join(HERE, name + ".h264")
        with open(path, "wb") as f:
            f.write(data)
        out = subprocess.run([os.path.join(ROOT, "oracle", "_ref", "ref_decode"), path], check=True,
                             capture_output=True, text=True, errors="replace").stdout
        pics = [json.loads(line[len("REFPIC "):]) for line in out.splitlines() if line.startswith("REFPIC ")]
        if len(pics) != n_pictures:
            sys.exit(f"the reference output {len(pics)} pictures for {src}, expected {n_pictures}")
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump({"source": src, "pictures": pics}, f, indent=1)
        print(f"{path}: {len(data)} bytes, {len(pics)} output pictures")


if __name__ == "__main__":
    main()
