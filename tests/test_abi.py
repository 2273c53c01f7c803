"""CPU tests of the product library's boundary: the C ABI loads and exports
every declared symbol, the host-side plan equals the oracle's, argument errors
map to the reference's exception kinds, and the seeded input generator."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import ROOT


def declared_functions(header: Path):
    text = header.read_text()
    return sorted(set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(bcrt_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol(pkg):
    lib = pkg.load_library()
    names = declared_functions(ROOT / "include" / "bytecrt.h")
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(pkg.bytecrt.EXPORTED)


def test_fherma_cpp_symbols_exported(pkg):
    so = pkg.bytecrt.LIB_PATH
    out = subprocess.run(["nm", "-D", "--defined-only", "-C", str(so)], capture_output=True, text=True,
                         check=True).stdout
    for sig in ("fherma_init(fherma::Point const&)", "fherma_run(void*, fherma::Inputs const&)",
                "fherma_free(void*)"):
        assert sig in out, sig


@pytest.mark.parametrize("N,W,seed", [(65536, 1747, 1), (32768, 868, 2), (16384, 436, 3), (8192, 220, 4),
                                      (4096, 100, 5)])
def test_host_plan_matches_oracle(pkg, orc, N, W, seed):
    q = pkg.datagen.modulus(W, seed)
    mine = pkg.bytecrt.plan_host(N, W, q)
    ref = orc.Plan(N, W, q)
    assert (mine["P"], mine["product_bits"]) == (ref.P, ref.product_bits)
    np.testing.assert_array_equal(mine["primes"], ref.primes)
    np.testing.assert_array_equal(mine["psi"], ref.psi)
    np.testing.assert_array_equal(mine["crt_inverse"], ref.crt_inv)
    np.testing.assert_array_equal(mine["weights"], ref.weights())
    assert mine["Kcrt"] == ref.kpad  # 4(P+1) padded to 32, as the reference's Kpad


def test_near_power_gap(pkg):
    """The CRT takes its near-power reduction exactly for q = 2^W - gap,
    0 < gap < 2^32 (the reference's only modulus form)."""
    gap_of = pkg.bytecrt.near_power_gap
    for W in (96, 220, 868, 1747):
        for gap in (1, 2, 65535, (1 << 31) + 1, (1 << 32) - 1):
            assert gap_of(W, (1 << W) - gap) == gap
        assert gap_of(W, (1 << W) - (1 << 32)) == 0
        assert gap_of(W, (1 << W) - (1 << 32) - 1) == 0
        assert gap_of(W, pkg.datagen.modulus(W, 3)) == 0  # random: top 32 bits not all ones
        assert gap_of(W, (1 << (W - 1)) + 1) == 0
    assert gap_of(64, (1 << 64) - 1) == 0  # two limbs: the generic steps


def test_serving_geometry_prime_counts(pkg):
    # N 8192..65536 with W 220/436/868/1747 -> 15/29/57/114 primes.
    for N, W, P in ((8192, 220, 15), (16384, 436, 29), (32768, 868, 57), (65536, 1747, 114)):
        assert pkg.bytecrt.plan_host(N, W, pkg.datagen.modulus(W, 0))["P"] == P


@pytest.mark.parametrize("N,W,q_bits,msg", [
    (3000, 220, 220, "power of two"),
    (131072, 220, 220, "power of two"),
    (8192, 220, 219, "bit_length"),
    (8192, 40, 40, "W"),
])
def test_host_plan_rejects_bad_geometry(pkg, N, W, q_bits, msg):
    q = (1 << (q_bits - 1)) | 1
    with pytest.raises(ValueError, match=msg):
        pkg.bytecrt.plan_host(N, W, q)


def test_fherma_init_error_kinds(pkg):
    bc = pkg.bytecrt
    L = 7
    q = bc.Tensor(shape=[L], data=bc.limbs_of((1 << 220) - 1, L))
    with pytest.raises(RuntimeError, match="B=1..128"):
        bc.fherma_init(bc.Point(B=0, N=8192, W=220, L=L, q=q))
    with pytest.raises(RuntimeError, match="B=1..128"):
        bc.fherma_init(bc.Point(B=129, N=8192, W=220, L=L, q=q))
    with pytest.raises(ValueError, match="limb count"):
        bc.fherma_init(bc.Point(B=1, N=8192, W=220, L=L, q=bc.Tensor(shape=[3], data=np.zeros(3, np.uint32))))


def test_datagen(pkg, orc):
    dg = pkg.datagen
    for W in (100, 220, 868):
        q = dg.modulus(W, 3)
        assert q.bit_length() == W and q & 1
        assert dg.modulus(W, 3) == q and dg.modulus(W, 4) != q
        x = dg.operands(1000, q, 5, edge_fraction=0.1, edges_first=True)
        vals = orc.rows_to_ints(x)
        assert all(0 <= v < q for v in vals)
        assert vals[:5] == [0, 1, q - 1, q // 2, q - q // 2]
        np.testing.assert_array_equal(x, dg.operands(1000, q, 5, edge_fraction=0.1, edges_first=True))


def test_probe_library_exports_every_declared_symbol(pkg):
    lib = pkg.probes.load_library()
    text = (ROOT / "include" / "bytecrt_probes.h").read_text()
    names = sorted(set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(bcrtp_\w+)\s*\(", text, flags=re.M)))
    assert names and set(names) == set(pkg.probes.EXPORTED)
    for n in names:
        assert hasattr(lib, n), n
