"""Writes h264_video_decoder_demo_amd/csrc/cabac_init.inc, the CABAC context
initialisation values (m, n) of ctxIdx 0..459 (the standard's Tables 9-12 to 9-24),
taken from the reference's own CH264Cabac::get_m_and_n through oracle/cabac_init_dump
(`make -C oracle cabac-init`), so that the table the front end uses is the one the
reference decodes with.  Run once where the reference is present:

    python3 oracle/make_cabac_init.py
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OUT = os.path.join(ROOT, "h264_video_decoder_demo_amd", "csrc", "cabac_init.inc")


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "cabac-init"], check=True, capture_output=True)
    text = subprocess.run([os.path.join(ROOT, "oracle", "_ref", "cabac_init_dump")], check=True, capture_output=True,
                          text=True).stdout
    rows = []
    for line in text.splitlines():
        f = line.split()
        ctx, vals = int(f[0]), f[1:]
        pairs = []
        for k in range(4):
            m, n = vals[2 * k], vals[2 * k + 1]
            pairs.append((0, 0) if m == "x" else (int(m), int(n)))
        assert ctx == len(rows)
        rows.append(pairs)
    with open(OUT, "w") as f:
        f.write("// CABAC context initialisation (9.3.1.1): m and n of ctxIdx 0..459 -- the standard's\n"
                "// Tables 9-12 to 9-24 -- for I slices, then cabac_init_idc 0, 1, 2 of P and B slices\n"
                "// (ctxIdx 276, the end_of_slice / terminate bin, has none).  Data, written by\n"
                "// oracle/make_cabac_init.py from the values the reference decodes with\n"
                "// (H264Cabac.cpp:37-1046).\n")
        f.write("static const int8_t kCabacInit[460][4][2] = {\n")
        for ctx, pairs in enumerate(rows):
            body = ", ".join("{%d, %d}" % (m, n) for m, n in pairs)
            f.write("    {" + body + "},  // " + str(ctx) + "\n")
        f.write("};\n")
    print(f"wrote {OUT}: {len(rows)} contexts")


if __name__ == "__main__":
    main()
