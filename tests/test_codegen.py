"""Checks on the gfx950 assembly that hipcc makes of the kernels (no GPU needed).

hipcc fuses Clip1(x >> n) of two values into v_ashr_pk_u8_i32 / v_ashr_pk_i8_i32 and
ORs further bytes onto the destination, whose upper half then keeps the register's
previous contents on gfx950: the luma interpolation came out wrong in columns 2-3
that way.  The kernels pack bytes with v_perm instead; this test keeps the fused
instructions out of every kernel.
"""
import os
import subprocess

import pytest

from h264_video_decoder_demo_amd import build as hbuild

FORBIDDEN = ("v_ashr_pk_u8_i32", "v_ashr_pk_i8_i32")


@pytest.fixture(scope="module")
def device_asm(tmp_path_factory):
    try:
        hipcc = hbuild.hipcc()
    except RuntimeError:
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("asm") / "kernels.s"
    flags = [f for f in hbuild.HIPCC_FLAGS if f not in ("-shared", "-fPIC")]
    cmd = [hipcc, *flags, "--cuda-device-only", "-S", "-I", os.path.join(hbuild.ROOT, "include"), "-I", hbuild.CSRC,
           os.path.join(hbuild.CSRC, "kernels.hip"), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return out.read_text()


def test_no_fused_shift_pack_instructions(device_asm):
    found = [ins for ins in FORBIDDEN if ins in device_asm]
    assert not found, f"kernels contain {found}"


def test_kernels_use_no_scratch(device_asm):
    """Every kernel runs without private (scratch) memory: spills would cost HBM traffic."""
    sizes = [int(line.split(":")[1]) for line in device_asm.splitlines()
             if line.strip().startswith(".private_segment_fixed_size:")]
    assert sizes and all(s == 0 for s in sizes), sizes


def test_kernels_use_no_flat_memory_instructions(device_asm):
    """Device memory is reached through global-address-space pointers: flat_* loads and
    stores would also count on the LDS counter, so every LDS wait would wait for them."""
    found = sorted({line.split()[0] for line in device_asm.splitlines()
                    if line.strip().startswith(("flat_load", "flat_store", "flat_atomic"))})
    assert not found, f"kernels contain {found}"
