"""GPU checks of the measurement probes (include/bytecrt_probes.h): the HBM
copy kernel copies exactly, the three butterfly forms (FP64 as the
product kernels compute, 32-bit Shoup, 32-bit Montgomery) produce the same
residues, equal to a Python-integer model of the same rounds, and the
MFMA/VALU overlap probe computes the same words alone and together."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

MASK64 = (1 << 64) - 1


def start_value(tid: int, k: int, p: int) -> int:
    z = (tid * 16 + k + 0x9E3779B97F4A7C15) & MASK64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return (z ^ (z >> 31)) % p


def model(tid: int, p: int, tw, rounds: int):
    v = [start_value(tid, k, p) for k in range(16)]
    for _ in range(rounds):
        for lvl in range(4):
            half = 8 >> lvl
            for g in range(1 << lvl):
                w = tw[(1 << lvl) - 1 + g]
                for x in range(half):
                    i0 = g * 2 * half + x
                    t = v[i0 + half] * w % p
                    v[i0], v[i0 + half] = (v[i0] + t) % p, (v[i0] - t) % p
    return v


def test_butterfly_forms_agree_with_model(pkg):
    pl = pkg.bytecrt.plan_host(65536, 1747, pkg.datagen.modulus(1747, 2026))
    for j in (0, 57, 113):  # primes near 2^31 and the smallest of the plan
        p, psi = int(pl["primes"][j]), int(pl["psi"][j])
        tw = [pow(psi, k, p) for k in range(1, 16)]
        threads, rounds = 512, 3
        outs = []
        # the five-operation forms (3, 4; primes of N = 65,536 have s >= 17)
        # also at four waves per SIMD
        for form, lds in ((0, 0), (1, 0), (2, 0), (3, 0), (4, 0), (4, pkg.probes.LDS_FOUR_WAVES), (5, 0)):
            out = torch.empty(threads * 16, dtype=torch.int32, device="cuda")
            pkg.probes.butterflies(form, out, threads, p, tw, rounds, lds_bytes=lds)
            torch.cuda.synchronize()
            outs.append(out.cpu().numpy().view("uint32"))
        for o in outs[1:]:
            assert (o == outs[0]).all()
        for tid in (0, 1, 255, 511):
            assert [int(x) for x in outs[0][16 * tid:16 * tid + 16]] == model(tid, p, tw, rounds), (j, tid)


def test_hbm_copy_probe_exact(pkg):
    n = 1 << 24
    src = torch.randint(-(1 << 31), (1 << 31) - 1, (n,), dtype=torch.int32, device="cuda")
    dst = torch.zeros_like(src)
    pkg.probes.hbm_stream(0, src, dst, n * 4)
    torch.cuda.synchronize()
    assert torch.equal(src, dst)
    pkg.probes.hbm_stream(2, None, dst, n * 4)
    pkg.probes.hbm_stream(1, src, dst, n * 4)
    torch.cuda.synchronize()


def test_issue_overlap_probe(pkg):
    pr = pkg.probes
    words = int(pr.load_library().bcrtp_issue_overlap_words())
    assert words > 0 and words % 512 == 0

    def run(mode, kind, mi, vi):
        out = torch.zeros(words, dtype=torch.int32, device="cuda")
        pr.issue_overlap(mode, kind, mi, vi, out)
        torch.cuda.synchronize()
        return out.cpu().numpy().reshape(-1, 512)

    alone_m, alone_v, both = run(1, 2, 50, 60), run(2, 2, 50, 60), run(3, 2, 50, 60)
    assert (both[:, :256] == alone_m[:, :256]).all() and (both[:, 256:] == alone_v[:, 256:]).all()
    assert not alone_m[:, 256:].any() and not alone_v[:, :256].any()
    # int kind: x <- (x mod 2^24) * 0x9E3779 + 12345 mod 2^32, eight chains from lane + k
    x = (np.arange(256, dtype=np.uint64) % 64)[:, None] + np.arange(8, dtype=np.uint64)
    for _ in range(60):
        x = ((x & 0xFFFFFF) * 0x9E3779 + 12345) & 0xFFFFFFFF
    want = (x.sum(axis=1) & 0xFFFFFFFF).astype(np.uint32).view(np.int32)
    assert (alone_v[:, 256:] == want[None, :]).all()
    # MFMA waves: the sums are linear in the iteration count (mod 2^32)
    twice = run(1, 2, 100, 60)
    assert (twice[:, :256] == (alone_m[:, :256].astype(np.int64) * 2).astype(np.int32)).all()
    for kind in (0, 1):
        a, b = run(2, kind, 1, 40), run(3, kind, 1, 40)
        assert (a[:, 256:] == b[:, 256:]).all()
    with pytest.raises(RuntimeError):
        pr.issue_overlap(4, 0, 1, 1, torch.zeros(words, dtype=torch.int32, device="cuda"))
