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
  tests/golden/gop121_idc0_12.h264       the same 12 pictures with every slice header's
      disable_deblocking_filter_idc rewritten from 2 to 0 (deblocking across the slice
      edges, which the reference filters with the current MB's data on both sides); the
      rest of each slice is moved by the two bits the shorter code saves
  tests/golden/gop121_poc1_12.h264       gop121's 12 pictures with pic_order_cnt_type 1
      (with_poc_type1 below: SPS cycle of two reference-frame offsets, an offset for
      non-reference pictures, delta_pic_order_cnt[0] in every slice header in place of
      pic_order_cnt_lsb)
  tests/golden/gop121_gap_12.h264        gop121's 12 pictures with frame_num raised by 2 from
      the 6th picture on (with_frame_num_gap below): a gap in frame_num, which the
      reference passes over without inserting frames
  tests/golden/heavyhand_cavlc_sl_8.h264  the first 8 pictures of the CAVLC clip with scaling
      matrices written into its SPS and PPS (with_scaling_matrices below: explicit lists,
      lists left to fall-back rules A and B, lists sent as "use the default", and PPS
      lists the reference takes from the SPS instead)
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
STREAMS = [("HeavyHand_1080p.B_frames_4.no_cabac.no_tff.h264", 15, "heavyhand_cavlc_15", None),
           ("HeavyHand_1080p.B_frames.cabac.h264", 10, "heavyhand_cabac_10", None),
           ("gop121.naluCnt453.h264", 12, "gop121_12", None),
           ("gop121.naluCnt453.h264", 12, "gop121_idc0_12", 0),
           ("HeavyHand_1080p.B_frames_4.no_cabac.no_tff.h264", 8, "heavyhand_cavlc_sl_8", "scaling"),
           ("gop121.naluCnt453.h264", 12, "gop121_poc1_12", "poc1"),
           ("gop121.naluCnt453.h264", 12, "gop121_gap_12", "gap")]


class _Bits:
    """Reads an RBSP bit by bit (7.2: u(n), ue(v), se(v)) and copies what it read."""

    def __init__(self, rbsp: bytes):
        self.bits = "".join(f"{b:08b}" for b in rbsp)
        self.pos = 0

    def u(self, n):
        v = int(self.bits[self.pos:self.pos + n] or "0", 2)
        self.pos += n
        return v

    def ue(self):
        z = 0
        while self.bits[self.pos + z] == "0":
            z += 1
        self.pos += z + 1
        return (1 << z) - 1 + self.u(z)

    def se(self):
        k = self.ue()
        return (k + 1) // 2 if k & 1 else -(k // 2)


def _unescape(nal: bytes) -> bytes:
    out, zeros = bytearray(), 0
    for b in nal:
        if zeros >= 2 and b == 3:
            zeros = 0
            continue
        out.append(b)
        zeros = zeros + 1 if b == 0 else 0
    return bytes(out)


def _escape(rbsp: bytes) -> bytes:
    out, zeros = bytearray(), 0
    for b in rbsp:
        if zeros >= 2 and b <= 3:
            out.append(3)
            zeros = 0
        out.append(b)
        zeros = zeros + 1 if b == 0 else 0
    return bytes(out)


def _ue_bits(v: int) -> str:
    b = bin(v + 1)[2:]
    return "0" * (len(b) - 1) + b


def _se_bits(v: int) -> str:
    return _ue_bits(2 * v - 1 if v > 0 else -2 * v)


def _list_bits(values, stop_at=None, default=False) -> str:
    """scaling_list() syntax for `values` (scan order): delta_scale per entry; `default`
    sends the first delta that makes nextScale 0 (useDefaultScalingMatrixFlag), `stop_at`
    ends the list there (the remaining entries repeat the last value)."""
    if default:
        return _se_bits(-8)
    bits, last = "", 8
    for j, v in enumerate(values):
        if j == stop_at:
            return bits + _se_bits(((0 - last + 128) % 256) - 128)
        bits += _se_bits(((v - last + 128) % 256) - 128)
        last = v
    return bits


# (present, bits) of the 8 scaling lists sent by the SPS and by the PPS
SPS_LISTS = [(1, _list_bits([8 + 2 * j for j in range(16)])),        # 0 Intra Y: a ramp
             (0, ""),                                                 # 1: rule A, copy of 0
             (1, _list_bits(None, default=True)),                     # 2: Default_4x4_Intra
             (0, ""),                                                 # 3: rule A, Default_4x4_Inter
             (1, _list_bits([12, 20] + [0] * 14, stop_at=2)),         # 4: 12, 20, 20, ...
             (0, ""),                                                 # 5: copy of 4
             (1, _list_bits([16 + j % 7 for j in range(64)])),        # 6 Intra Y 8x8
             (1, _list_bits(None, default=True))]                     # 7: Default_8x8_Inter
PPS_LISTS = [(1, _list_bits([20] * 16)),                             # 0: taken from the SPS's list 0
             (0, ""), (0, ""),                                        # 1, 2: copies (rule B)
             (1, _list_bits(None, default=True)),                     # 3: Default_4x4_Inter
             (1, _list_bits([30] * 16)),                              # 4: taken from the SPS's list 4
             (0, ""),                                                 # 5: copy of 4
             (0, ""),                                                 # 6: rule B, the SPS-level list
             (1, _list_bits([40] * 64))]                              # 7: the SPS's parsed list 7 (all 8)


def with_scaling_matrices(data: bytes) -> bytes:
    """The stream with seq_scaling_matrix_present_flag and pic_scaling_matrix_present_flag
    set and the lists above written after them (7.3.2.1.1, 7.3.2.2), for a High-profile
    4:2:0 SPS and a PPS with transform_8x8_mode_flag; everything else is kept."""
    starts, i = [], 0
    while (j := data.find(b"\x00\x00\x01", i)) >= 0:
        starts.append(j)
        i = j + 3
    out = bytearray(data[:starts[0]])
    for k, st in enumerate(starts):
        end = starts[k + 1] if k + 1 < len(starts) else len(data)
        nal = data[st + 3:end]
        while k + 1 < len(starts) and nal.endswith(b"\x00"):
            nal = nal[:-1]
        trail = data[st + 3 + len(nal):end]
        hdr, typ = nal[0], nal[0] & 31
        if typ in (7, 8):
            b = _Bits(_unescape(nal[1:]))
            if typ == 7:
                assert b.u(8) == 100, "High profile"
                b.u(16); b.ue()
                assert b.ue() == 1, "4:2:0"
                b.ue(); b.ue(); b.u(1)
                lists = SPS_LISTS
            else:
                b.ue(); b.ue(); b.u(2)
                assert b.ue() == 0
                b.ue(); b.ue(); b.u(3); b.se(); b.se(); b.se(); b.u(3)
                assert b.u(1) == 1, "transform_8x8_mode_flag"
                lists = PPS_LISTS
            at = b.pos
            assert b.u(1) == 0, "no scaling matrix yet"
            body = b.bits.rstrip("0")
            pad = (len(b.bits) - len(body)) // 8
            bits = b.bits[:at] + "1" + "".join(f"{p}{x}" for p, x in lists) + body[b.pos:-1] + "1"
            bits += "0" * (-len(bits) % 8 + 8 * pad)
            nal = bytes([hdr]) + _escape(bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8)))
        out += data[st:st + 3] + nal + trail
    return bytes(out)


# pic_order_cnt_type 1 fields written into the SPS: delta_pic_order_always_zero_flag 0,
# offset_for_non_ref_pic -1, offset_for_top_to_bottom_field 0, a cycle of two
# offset_for_ref_frame (3, 1); the k-th picture's slices send delta_pic_order_cnt[0] k % 2
POC1_SPS = "0" + _se_bits(-1) + _se_bits(0) + _ue_bits(2) + _se_bits(3) + _se_bits(1)


def _nals(data: bytes):
    """(start code offset, NAL bytes, bytes after the NAL up to the next start code)."""
    starts, i = [], 0
    while (j := data.find(b"\x00\x00\x01", i)) >= 0:
        starts.append(j)
        i = j + 3
    for k, st in enumerate(starts):
        end = starts[k + 1] if k + 1 < len(starts) else len(data)
        nal = data[st + 3:end]
        while k + 1 < len(starts) and nal.endswith(b"\x00"):
            nal = nal[:-1]
        yield st, nal, data[st + 3 + len(nal):end]


def _repack(hdr: int, bits: str, pad: int) -> bytes:
    """A NAL from its header byte and its RBSP bits up to the stop bit (exclusive)."""
    bits += "1"
    bits += "0" * (-len(bits) % 8 + 8 * pad)
    return bytes([hdr]) + _escape(bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8)))


def with_poc_type1(data: bytes) -> bytes:
    """gop121 (Baseline, POC type 0, CAVLC I / P slices) with pic_order_cnt_type 1: the
    SPS's log2_max_pic_order_cnt_lsb_minus4 becomes the type-1 fields above, and each
    slice's pic_order_cnt_lsb becomes delta_pic_order_cnt[0] (7.3.2.1.1, 7.3.3)."""
    out, sps, pps, pic = bytearray(data[:data.find(b"\x00\x00\x01")]), {}, {}, -1
    for st, nal, trail in _nals(data):
        hdr, typ = nal[0], nal[0] & 31
        b = _Bits(_unescape(nal[1:]))
        body = b.bits.rstrip("0")
        pad = (len(b.bits) - len(body)) // 8
        if typ == 7:
            assert b.u(8) == 66, "Baseline"
            b.u(16); b.ue()
            sps["log2_fn"] = b.ue() + 4
            at = b.pos
            assert b.ue() == 0, "POC type 0"
            sps["log2_poc"] = b.ue() + 4
            nal = _repack(hdr, body[:at] + _ue_bits(1) + POC1_SPS + body[b.pos:-1], pad)
        elif typ == 8:
            b.ue(); b.ue(); b.u(1)
            assert b.u(1) == 0, "no bottom_field_pic_order_in_frame_present_flag"
        elif typ in (1, 5):
            if b.bits[b.pos] == "1":  # first_mb_in_slice 0: a new picture
                pic += 1
            b.ue(); b.ue(); b.ue(); b.u(sps["log2_fn"])
            if typ == 5:
                b.ue()
            at = b.pos
            b.u(sps["log2_poc"])
            nal = _repack(hdr, body[:at] + _se_bits(pic % 2) + body[b.pos:-1], pad)
        out += data[st:st + 3] + nal + trail
    return bytes(out)


def with_frame_num_gap(data: bytes, first: int, by: int) -> bytes:
    """gop121 with frame_num + `by` (modulo MaxFrameNum) in the slices of its pictures from
    index `first` on: the field keeps its width, so only those bits change (7.3.3)."""
    out, sps, pic = bytearray(data[:data.find(b"\x00\x00\x01")]), {}, -1
    for st, nal, trail in _nals(data):
        hdr, typ = nal[0], nal[0] & 31
        b = _Bits(_unescape(nal[1:]))
        if typ == 7:
            b.u(24); b.ue()
            sps["log2_fn"] = b.ue() + 4
        elif typ in (1, 5):
            if b.bits[b.pos] == "1":
                pic += 1
            if pic >= first:
                b.ue(); b.ue(); b.ue()
                at, n = b.pos, sps["log2_fn"]
                fn = (b.u(n) + by) % (1 << n)
                body = b.bits.rstrip("0")
                pad = (len(b.bits) - len(body)) // 8
                nal = _repack(hdr, body[:at] + format(fn, f"0{n}b") + body[at + n:-1], pad)
        out += data[st:st + 3] + nal + trail
    return bytes(out)


def with_deblocking_idc(data: bytes, idc: int) -> bytes:
    """The stream with disable_deblocking_filter_idc of every slice set to idc, for the
    syntax gop121 uses (CAVLC I / P slices, POC type 0, no weighted prediction): SPS and
    PPS fields are read as far as the slice header needs them (7.3.2.1, 7.3.2.2), the
    slice header up to the field (7.3.3), and the slice's bits after it are kept."""
    starts, i = [], 0
    while (j := data.find(b"\x00\x00\x01", i)) >= 0:
        starts.append(j)
        i = j + 3
    out, sps, pps = bytearray(data[:starts[0]]), {}, {}
    for k, st in enumerate(starts):
        end = starts[k + 1] if k + 1 < len(starts) else len(data)
        nal = data[st + 3:end]
        while k + 1 < len(starts) and nal.endswith(b"\x00"):  # the next start code's leading zero
            nal = nal[:-1]
        trail = data[st + 3 + len(nal):end]
        hdr, typ = nal[0], nal[0] & 31
        b = _Bits(_unescape(nal[1:]))
        if typ == 7:
            prof = b.u(8); b.u(16); b.ue()
            assert prof in (66, 77), "the rewrite reads Baseline / Main parameter sets"
            sps = {"log2_fn": b.ue() + 4, "poc_type": b.ue()}
            assert sps["poc_type"] == 0
            sps["log2_poc"] = b.ue() + 4
        elif typ == 8:
            b.ue(); b.ue()
            assert b.u(1) == 0, "CAVLC"
            pps["poc_present"] = b.u(1)
            assert b.ue() == 0, "one slice group"
            pps["nref"] = (b.ue() + 1, b.ue() + 1)
            assert b.u(1) == 0 and b.u(2) == 0, "no weighted prediction"
            b.se(); b.se(); b.se()
            pps["dbk_present"] = b.u(1)
            b.u(1)
            pps["redundant"] = b.u(1)
        elif typ in (1, 5) and pps.get("dbk_present"):
            b.ue()
            st_type = b.ue() % 5
            assert st_type in (0, 2), "I and P slices"
            b.ue(); b.u(sps["log2_fn"])
            if typ == 5:
                b.ue()
            b.u(sps["log2_poc"])
            if pps["poc_present"]:
                b.se()
            if pps["redundant"]:
                b.ue()
            if st_type == 0:
                if b.u(1):
                    b.ue()
                if b.u(1):  # ref_pic_list_modification_flag_l0
                    while (op := b.ue()) != 3:
                        b.ue()
            if hdr >> 5:
                if typ == 5:
                    b.u(2)
                elif b.u(1):
                    while (op := b.ue()) != 0:
                        if op in (1, 3):
                            b.ue()
                        if op in (2, 3, 6):
                            b.ue()
                        if op == 4:
                            b.ue()
            b.se()  # slice_qp_delta
            at = b.pos
            old = b.ue()
            body = b.bits.rstrip("0")
            pad = (len(b.bits) - len(body)) // 8  # whole zero bytes after the stop bit's byte
            rest = body[b.pos:-1]  # the bits before rbsp_stop_one_bit
            bits = b.bits[:at] + _ue_bits(idc if old != 1 else old) + rest + "1"
            bits += "0" * (-len(bits) % 8 + 8 * pad)
            nal = bytes([hdr]) + _escape(bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8)))
        out += data[st:st + 3] + nal + trail
    return bytes(out)


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
    for src, n_pictures, name, idc in STREAMS:
        with open(os.path.join(DATA, src), "rb") as f:
            data = prefix(f.read(), n_pictures)
        if idc == "scaling":
            data = with_scaling_matrices(data)
        elif idc == "poc1":
            data = with_poc_type1(data)
        elif idc == "gap":
            data = with_frame_num_gap(data, 5, 2)
        elif idc is not None:
            data = with_deblocking_idc(data, idc)
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
