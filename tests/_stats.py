"""Comparison of DistributionStats with the oracle's (tests only).

The oracle sums sequentially in f64, in linear index order, as the reference
does (src/metrics.rs:215-226).  The library's default order (exact device
sums rounded once) differs from that only by the rounding of the sequential
sums:
  * count, maximum and the five quantiles are order statistics or exact
    counts: equal;
  * the mean is equal wherever the reference's running sum is itself exact:
    every value is an f32 of at least 2 voxels times the spacing (a diameter
    of at least 2), so a multiple of g, the ulp of that smallest value, and a
    running sum below 2^53 g takes no rounding.  Elsewhere it is within the
    sequential sum's worst case, (n - 1) 2^-53 relative;
  * the SD inherits the rounding of the sequential sum of squares, worst case
    (n - 1) 2^-53 relative, amplified by the cancellation in
    n*sum(x^2) - sum(x)^2 by (1 + 1/CV^2), halved by the square root.
    Measured, that rounding is far from a random walk: it moves the
    reference's SD by 2.6e-12 to 1.3e-11 relative on volumes of 10^5 to 10^6
    values and by 2.9e-10 on the 512^3 workload, so 1e-12 does not hold for
    the SD in this order.
With the reference's order (bm.StatsOrder.Sequential) everything is equal."""
import math

REL = 1e-12
QUANTILE_FIELDS = ("median", "percentile_10", "percentile_25", "percentile_75",
                   "percentile_90", "interquartile_range")
FIELDS = ("mean", "standard_deviation", "maximum") + QUANTILE_FIELDS + (
    "coefficient_of_variation", "count")


def sequential_bound(n):
    """Worst-case relative error of a sequential f64 sum of n positive terms."""
    return max(n - 1, 0) * 2.0 ** -53


def exact_running_sum(n, maximum, scale):
    """True when the reference's running sum of n values, none above maximum
    and all multiples of the ulp of the f32 value 2*scale, stays below 2^53
    of those ulps (no rounding at all)."""
    g = 2.0 ** (math.floor(math.log2(2.0 * scale)) - 23)
    return n * maximum < 2.0 ** 53 * g


def _get(s, f):
    return s[f] if isinstance(s, dict) else getattr(s, f)


def _eq(a, b):
    return a == b or (isinstance(a, float) and math.isnan(a) and math.isnan(b))


def relative_differences(got, ref):
    """{field: |got - ref| / |ref|} of mean, SD and CV."""
    out = {}
    for f in ("mean", "standard_deviation", "coefficient_of_variation"):
        g, r = _get(got, f), _get(ref, f)
        out[f] = abs(g - r) / abs(r) if r else abs(g - r)
    return out


def check_stats(got, ref, sequential=False, scale=None):
    """got (library) against ref (oracle): order statistics and counts equal;
    with sequential=True everything equal, else mean and SD within the
    reduction-order bounds above (the mean equal where the reference's sum is
    exact, when the spacing `scale` is given)."""
    assert _get(got, "count") == _get(ref, "count")
    assert _get(got, "maximum") == _get(ref, "maximum")
    for f in QUANTILE_FIELDS:
        assert _eq(_get(got, f), _get(ref, f)), (f, _get(got, f), _get(ref, f))
    if sequential:
        for f in ("mean", "standard_deviation", "coefficient_of_variation"):
            assert _get(got, f) == _get(ref, f), (f, _get(got, f), _get(ref, f))
        return
    n = _get(ref, "count")
    cv = _get(ref, "coefficient_of_variation")
    d = relative_differences(got, ref)
    if scale is not None and exact_running_sum(n, _get(ref, "maximum"), scale):
        assert _get(got, "mean") == _get(ref, "mean"), ("mean", d)
    mtol = max(REL, sequential_bound(n))
    stol = max(REL, sequential_bound(n) * (1.0 + 1.0 / cv ** 2) / 2.0) if cv else REL
    assert d["mean"] <= mtol, ("mean", d, mtol)
    assert d["standard_deviation"] <= stol, ("sd", d, stol)
    assert d["coefficient_of_variation"] <= stol + mtol, ("cv", d, stol)
