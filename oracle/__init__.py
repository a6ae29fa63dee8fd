"""The FP64 oracle: a serial C restatement of the reference's per-step path.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py may import this package; the product path
(bluebottle_amd and its libraries) never does.
"""
