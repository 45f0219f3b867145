"""TEST INFRASTRUCTURE ONLY: the CPU oracle of the reconstruction engine.

``oracle/refcpu.c`` restates the reference decoder's picture reconstruction
(jfu222/h264_video_decoder_demo) in plain C; ``oracle/ref_harness.cpp`` drives
the reference's own sources (built into ``oracle/_ref/`` when /root/reference is
present) to pin that restatement.  Only ``tests/``, ``__graft_entry__.smoke()``
and bench.py's CPU-baseline leg import this package, as the checker; the engine
never calls it.
"""
