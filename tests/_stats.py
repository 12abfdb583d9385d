"""Comparison of DistributionStats with the oracle's (tests only).

The oracle sums sequentially in f64, in linear index order, as the reference
does (src/metrics.rs:215-226).  The library's default order (exact device
sums rounded once) differs from that only by the rounding of the sequential
sums:
  * count, maximum and the five quantiles are order statistics or exact
    counts: equal;
  * the mean is equal wherever the sequential sum of the values is itself
    exact (every value is a multiple of the ulp of the smallest one, so the
    running sum stays exact below 2^53 of those ulps) and otherwise within
    the sequential sum's rounding;
  * the SD inherits the rounding of the sequential sum of squares, amplified
    by the cancellation in n*sum(x^2) - sum(x)^2 by (1 + 1/CV^2), halved by
    the square root.
REL is the 1e-12 relative tolerance of the statistics.  rounding_bound(n)
is the statistical size of n sequential roundings of relative size 2^-53
each (a random walk: sqrt(n) 2^-53), with a factor 8 of margin; the tests
take max(REL, that) so that the bound only matters beyond about 10^7 values.
With the sequential order (bm.StatsOrder.Sequential) everything is equal."""
import math

REL = 1e-12
QUANTILE_FIELDS = ("median", "percentile_10", "percentile_25", "percentile_75",
                   "percentile_90", "interquartile_range")
FIELDS = ("mean", "standard_deviation", "maximum") + QUANTILE_FIELDS + (
    "coefficient_of_variation", "count")


def rounding_bound(n):
    return 8.0 * math.sqrt(max(n, 1)) * 2.0 ** -53


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


def check_stats(got, ref, sequential=False):
    """got (library) against ref (oracle): order statistics and counts equal;
    with sequential=True everything equal, else mean/SD/CV within the
    reduction-order bounds above."""
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
    mtol = max(REL, rounding_bound(n))
    stol = max(REL, rounding_bound(n) * (1.0 + 1.0 / cv ** 2) / 2.0) if cv else REL
    assert d["mean"] <= mtol, ("mean", d, mtol)
    assert d["standard_deviation"] <= stol, ("sd", d, stol)
    assert d["coefficient_of_variation"] <= stol + mtol, ("cv", d, stol)
