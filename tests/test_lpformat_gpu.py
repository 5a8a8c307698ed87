"""An LP file read by the LP-format reader and solved on the GPU: the same
status, iteration count and objective as the oracle's restated IPM on the
parsed problem (the file carries 6 significant digits, so the parsed LP,
not the generator's, is the common input)."""
import pytest

import lpformat
import lpgen
import oracle


@pytest.mark.gpu
@pytest.mark.parametrize("seed,m,n", [(11, 40, 100), (12, 120, 300)])
def test_parsed_lp_solves_like_the_oracle(hip, seed, m, n, tmp_path):
    print(f"[case] lp-format seed {seed} {m}x{n}", flush=True)
    path = str(tmp_path / "p.lp")
    lpformat.write_lp(lpgen.random_lp(seed, m, n), path)
    lp, _ = lpformat.read_lp(path)
    assert lp.m == m and lp.n == n
    want = oracle.ipm_solve(lp)
    got = hip.ipm_solve(lp)
    assert got.status == want.status and got.iterations == want.iterations
    assert abs(got.objective_value - want.objective) <= 1e-8 * (1.0 + abs(want.objective))
