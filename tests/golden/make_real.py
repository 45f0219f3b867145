"""Writes the real-stream fixtures: a prefix of one of the reference's own sample
streams and the reference decoder's output for it.

  tests/golden/heavyhand_cavlc_15.h264   the first 15 pictures (whole NAL units) of
      data/HeavyHand_1080p.B_frames_4.no_cabac.no_tff.h264: High profile, CAVLC,
      1920x1088 frames, B pyramids, explicit (P) and implicit (B) weighted prediction,
      4 reference frames, reference list modification
  tests/golden/heavyhand_cavlc_15.json   per output picture of the reference decoder
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
SRC = "/root/reference/data/HeavyHand_1080p.B_frames_4.no_cabac.no_tff.h264"
N_PICTURES = 15
NAME = "heavyhand_cavlc_15"


def prefix(data: bytes, n_pictures: int) -> bytes:
    """The stream up to the end of its n-th coded slice NAL unit."""
    starts, i = [], 0
    while True:
        j = data.find(b"\x00\x00\x01", i)
        if j < 0:
            break
        starts.append(j)
        i = j + 3
    pics = 0
    for k, s in enumerate(starts):
        if data[s + 3] & 31 in (1, 5):
            pics += 1
            if pics == n_pictures:
                end = starts[k + 1] if k + 1 < len(starts) else len(data)
                return data[:end]
    raise ValueError("the stream has fewer pictures")


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "decode"], check=True, capture_output=True)
    with open(SRC, "rb") as f:
        data = prefix(f.read(), N_PICTURES)
    path = os.path.join(HERE, NAME + ".h264")
    with open(path, "wb") as f:
        f.write(data)
    out = subprocess.run([os.path.join(ROOT, "oracle", "_ref", "ref_decode"), path], check=True, capture_output=True,
                         text=True, errors="replace").stdout
    pics = [json.loads(line[len("REFPIC "):]) for line in out.splitlines() if line.startswith("REFPIC ")]
    if len(pics) != N_PICTURES:
        sys.exit(f"the reference output {len(pics)} pictures, expected {N_PICTURES}")
    with open(os.path.join(HERE, NAME + ".json"), "w") as f:
        json.dump({"source": os.path.basename(SRC), "pictures": pics}, f, indent=1)
    print(f"{path}: {len(data)} bytes, {len(pics)} output pictures")


if __name__ == "__main__":
    main()
