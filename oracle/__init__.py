"""CPU oracle of the explicit Euler residual loop: test infrastructure only.

Only the tests, the smoke check and the benchmark's CPU baseline use it, as the checker; the
product path (autan_amd) never imports it.
"""
