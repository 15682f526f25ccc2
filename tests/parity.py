"""Comparison helpers for GPU-vs-oracle parity (imported by the tests)."""
import numpy as np

REL_TOL = 1e-5  # screen positions, depths, lit colours: 1e-5 relative


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    same_inf = np.isinf(a) & np.isinf(b) & (np.sign(a) == np.sign(b))
    d = np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), 1e-30)
    d[both_nan | same_inf | (a == b)] = 0
    return d


def compare_frames(gpu, ora, exact=True):
    """Render pairs, survivor sets and draw order bit-exact; vertex fields bit-exact
    (exact=True) or within REL_TOL.  Returns a summary dict."""
    assert [p[:4] for p in gpu.pairs] == [p[:4] for p in ora.pairs], "render pairs / survivor counts differ"
    assert gpu.color.shape == ora.color.shape
    np.testing.assert_array_equal(gpu.triangle_ids, ora.triangle_ids)
    np.testing.assert_array_equal(gpu.indices, ora.indices)
    summary = {"pairs": len(gpu.pairs), "survivors": gpu.survivors, "max_rel": 0.0, "bit_mismatch": 0}
    gn, on = getattr(gpu, "normals", None), getattr(ora, "normals", None)
    assert (gn is None) == (on is None), "normal list present on one side only"
    names = ("color", "depth") if gn is None else ("color", "depth", "normals")
    for name in names:
        g, o = getattr(gpu, name), getattr(ora, name)
        mism = int(np.count_nonzero(g.view(np.uint32) != o.view(np.uint32)))
        e = rel_err(g, o)
        emax = float(e.max()) if e.size else 0.0   # a frame may draw nothing
        summary["bit_mismatch"] += mism
        summary["max_rel"] = max(summary["max_rel"], emax)
        if exact and mism:
            bad = np.argwhere(g.view(np.uint32) != o.view(np.uint32))[:8]
            detail = ["v%d f%d gpu=%r oracle=%r" % (i, j, float(g[i, j]), float(o[i, j])) for i, j in bad]
            raise AssertionError("%s: %d bitwise mismatches, max rel %.3g; first: %s" % (name, mism, emax, detail))
        else:
            assert emax <= REL_TOL, "%s: max rel %.3g" % (name, emax)
    return summary
