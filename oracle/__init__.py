"""CPU oracle for the per-time-step hot path of jaxley-pytorch's `jx.integrate`.

TEST INFRASTRUCTURE ONLY. Only `tests/`, `__graft_entry__.smoke()` and the
`cpu_baseline` leg of `bench.py` may import this package, and only as the checker
(or as the timed CPU baseline). The product (`jaxhip`) never imports it.
"""
