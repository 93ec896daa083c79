"""The C oracle against the reference's own outputs (golden fixtures written by
tests/golden/make_golden.py from the compiled reference)."""
import hashlib

import numpy as np
import pytest

from oracle import pyoracle as po


def _sha(b: bytes) -> str:
    return hashlib.sha256(b).hexdigest()


def test_golden_archives_bit_exact(golden):
    bad = []
    for c in golden["cases"]:
        a, nout, _ = po.compress(c["x"], c["eb"], c["r2r"], c["dict"], c["block"])
        if a != c["archive"] or nout != c["n_outlier"]:
            bad.append(c["name"])
    assert not bad, f"oracle archive differs from the reference for {bad}"


def test_golden_decompress_bit_exact(golden):
    bad = []
    for c in golden["cases"]:
        y = po.decompress(c["archive"], c["x"].shape, c["r2r"], c["value_range"], c["dict"], c["block"])
        if not np.array_equal(y.view(np.uint32), c["out"].view(np.uint32)):
            bad.append(c["name"])
    assert not bad, f"oracle decompression differs from the reference for {bad}"


def test_golden_covers_every_rule(golden):
    names = {c["name"] for c in golden["cases"]}
    shapes = {len(c["x"].shape) for c in golden["cases"]}
    assert shapes == {1, 2, 3}
    blocks = {c["block"] for c in golden["cases"]}
    assert blocks == {8, 16, 32, 64}
    # partial x-group in 2D (lost columns) and 3D (scalar path)
    assert any(c["x"].ndim == 2 and c["x"].shape[-1] % 8 for c in golden["cases"])
    assert any(c["x"].ndim == 3 and c["x"].shape[-1] % 8 for c in golden["cases"])
    for n in ("constant_2d", "all_outlier_3d", "ties_3d_scalar_group", "specials_3d", "dict_65536_u16_tree"):
        assert n in names


def test_golden_digest_u32_tree():
    """> 65536 tree nodes (u32 L/R arrays): input is integer-exact, so it is
    regenerated bit for bit and the oracle archive must hash like the reference's."""
    import json
    import os

    from tests.golden.cases import digest_cases

    man = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))
    ref = {d["name"]: d for d in man["digests"]}
    c = [c for c in digest_cases() if c["name"] == "u32_tree_1d"][0]
    d = ref["u32_tree_1d"]
    assert _sha(c["x"].tobytes()) == d["sha256_input"]
    assert d["node_count"] > 65536
    a, nout, _ = po.compress(c["x"], c["eb"], c["r2r"], c["dict"], c["block"])
    assert _sha(a) == d["sha256_archive"] and nout == d["n_outlier"]
    y = po.decompress(a, c["x"].shape, False, 0.0, c["dict"], c["block"])
    assert _sha(y.tobytes()) == d["sha256_out"]


@pytest.mark.parametrize("name", ["c1_sinusoids_2d_full", "grf_128_r2r_1e-4"])
def test_golden_digest_synthetic(name):
    """Full-size synthetic fields: valid where the generator reproduces the
    stored input bits (numpy transcendental functions may differ per host)."""
    import json
    import os

    from tests.golden.cases import digest_cases

    man = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))
    d = {x["name"]: x for x in man["digests"]}[name]
    c = [c for c in digest_cases() if c["name"] == name][0]
    if _sha(c["x"].tobytes()) != d["sha256_input"]:
        pytest.skip("generator does not reproduce the stored input bits on this host")
    a, nout, _ = po.compress(c["x"], c["eb"], c["r2r"], c["dict"], c["block"])
    assert _sha(a) == d["sha256_archive"] and nout == d["n_outlier"]
    rng = po.value_range(c["x"]) if c["r2r"] else 0.0
    y = po.decompress(a, c["x"].shape, c["r2r"], rng, c["dict"], c["block"])
    assert _sha(y.tobytes()) == d["sha256_out"]
