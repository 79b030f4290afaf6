"""Host logic of the serving loop that needs no GPU."""
import math

from magicube_amd import attention, serving


def test_percentile_matches_linear_interpolation():
    vals = sorted([5.0, 1.0, 3.0, 2.0, 4.0])
    assert serving._percentile(vals, 50) == 3.0
    assert serving._percentile(vals, 0) == 1.0 and serving._percentile(vals, 100) == 5.0
    assert math.isclose(serving._percentile(vals, 99), 4.96)


def test_working_set_counts_quantized_operands_and_output():
    # int8 q, k, v (L*64 each) + fp32 o (L*64*4)
    assert attention.working_set_bytes(1024, 8) == 1024 * 64 * 7
    # int4 q, k halve
    assert attention.working_set_bytes(1024, 4) == 1024 * 64 * 6
