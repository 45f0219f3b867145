"""The C oracle against the reference decoder's own output.

tests/golden/golden.json holds SHA-256 of every picture the reference's
reconstruction sources produced on the generated syntax (make_golden.py drives
them through oracle/ref_harness.cpp).  The oracle must reproduce each one, before
and after deblocking, and the generator must reproduce the syntax digests.
"""
import pytest

from helpers import make_generator, oracle_decode, plane_hashes

CPU_CASES = ["cif_ip", "cif_ip_seed2", "cif_ip_seed3", "cif_pcm_cqp", "high_small", "high_small_seed7",
             "high_implicit", "high_default_nodbk", "narrow", "flat", "high_1080p", "high_720p", "high_2160p",
             "mbaff_small", "mbaff_t8_pcm", "mbaff_tall", "mbaff_1080p", "high_small_cip", "high_odd_cip",
             "high_direct", "high_direct_odd", "high_direct_pcm"]
BIG = ("high_1080p", "high_720p", "high_2160p", "mbaff_1080p")


def test_golden_covers_cases(golden):
    assert sorted(golden) == sorted(CPU_CASES)


@pytest.mark.parametrize("name", CPU_CASES)
def test_generator_digest(golden, name):
    case = golden[name]
    gen = make_generator(case)
    assert f"{gen.digest:016x}" == case["digest"]
    assert gen.num_pictures == case["pictures"]
    assert (gen.width_mbs, gen.height_mbs) == (case["width_mbs"], case["height_mbs"])


@pytest.mark.parametrize("name", CPU_CASES)
def test_oracle_matches_reference(golden, name):
    case = golden[name]
    gen = make_generator(case)
    stages = (1, 2) if name not in BIG else (2,)
    for i, desc, out in oracle_decode(gen, stages):
        if 1 in out:
            assert plane_hashes(*out[1]) == case["pre"][i], f"{name} picture {i} before deblocking"
        assert plane_hashes(*out[2]) == case["pic"][i], f"{name} picture {i}"


@pytest.mark.parametrize("name", ["cif_ip", "high_small", "flat", "mbaff_small"])
def test_oracle_bgr24_matches_reference(golden, name):
    """YUV420P -> BGR24 output conversion against the reference's convertYuv420pToBgr24."""
    import hashlib
    from oracle.refcpu import yuv420p_to_bgr24
    case = golden[name]
    for i, desc, out in oracle_decode(make_generator(case), (2,)):
        bgr = yuv420p_to_bgr24(*out[2])
        assert hashlib.sha256(bgr.tobytes()).hexdigest() == case["bgr"][i], f"{name} picture {i}"
