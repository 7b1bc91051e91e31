"""The float path's wet/dry tests are the reference's double tests, exactly.

The reference compares depths with the double literal 1e-6 (h < 1e-6,
h > 1e-6, h <= 1e-6: flucCal_2ndOrder_jh_improved_Kernel.cu:44-72,
fluxCal_jh_modified_surface_Kernel.cu:96-139, timeControl_Kernel.cu:164-198,
friction_implicit_andUpdate_jh_Kernel.cu:22-146), so a float depth is widened
to double before the test.  The HIP float path (swe_physics.hpp: lt_hs,
le_hs, gt_hs) compares in float against fl(1e-6), the float nearest 1e-6,
which lies below 1e-6 with the next float above it; then for every float x

    x <  1e-6  <=>  x <= fl(1e-6)
    x <= 1e-6  <=>  x <= fl(1e-6)
    x >  1e-6  <=>  x >  fl(1e-6)

and the float tests decide as the double ones.  The new depth of the dry test
(h + dh, new_depth_dry) is formed in double: the sum of two floats whose
exponents differ by less than 29 is exact in double, so it is the real sum.
"""
import numpy as np

HS = 1e-6


def test_float_threshold_lies_below_the_double_one_with_the_next_float_above():
    f = np.float32(HS)
    assert float(f) < HS
    assert float(np.nextafter(f, np.float32(np.inf))) > HS


def test_float_tests_decide_as_double_tests_around_the_threshold():
    f = np.float32(HS)
    xs = [f]
    up = down = f
    for _ in range(2000):
        up = np.nextafter(up, np.float32(np.inf))
        down = np.nextafter(down, np.float32(0))
        xs += [up, down]
    xs += [np.float32(0), np.float32(-1e-6), np.float32(1.0), np.float32(-0.0), np.float32(1e-45)]
    for x in xs:
        d = float(x)
        assert (d < HS) == (x <= f)
        assert (d <= HS) == (x <= f)
        assert (d > HS) == (x > f)


def test_new_depth_is_exact_in_double():
    rng = np.random.Generator(np.random.PCG64(5))
    h = rng.uniform(0, 2e-6, 100000).astype(np.float32)
    dh = (rng.uniform(-2e-6, 2e-6, 100000) * rng.choice([1e-3, 1.0, 1e3], 100000)).astype(np.float32)
    s = h.astype(np.float64) + dh.astype(np.float64)
    # exact: the double sum minus each term gives the other term back
    assert np.array_equal(s - h.astype(np.float64), dh.astype(np.float64))
    # a sum the float addition rounds back below the threshold: fl(1e-6) plus
    # 3e-15 (under half an ulp) is 1e-6 + 5e-16 exactly, wet for the
    # reference, while the float sum is fl(1e-6) again
    f = np.float32(HS)
    d = np.float32(3e-15)
    assert f + d == f
    assert float(f) + float(d) > HS
