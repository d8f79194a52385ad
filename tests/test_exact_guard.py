"""The predicate that lets the BF16 local join skip the exact FP32 re-score
(kgraph_amd/csrc/kg_exact.hpp), compiled with g++ and checked on the CPU
against cases whose arithmetic is worked out here: the tiles are exact only
on integer rows BF16 holds, with 2 D max|x|^2 <= 2^24 and (except for the
angular metric) D (max - min)^2 <= 2^24."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "kgraph_amd", "csrc")

PROGRAM = r'''
#include <cstdio>
#include "kg_exact.hpp"
int main() {
    unsigned D; int metric, integer, lossless; double mx, lo, hi;
    while (std::scanf("%u %d %d %d %lf %lf %lf", &D, &metric, &integer, &lossless, &mx, &lo, &hi) == 7)
        std::printf("%d\n", kga::bf16_tiles_exact(D, metric, integer, lossless, mx, lo, hi) ? 1 : 0);
    return 0;
}
'''


@pytest.fixture(scope="module")
def predicate(tmp_path_factory):
    d = tmp_path_factory.mktemp("exact")
    src, exe = d / "p.cpp", d / "p"
    src.write_text(PROGRAM)
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + CSRC, str(src), "-o", str(exe)], check=True)

    def run(cases):
        inp = "".join("%d %d %d %d %.17g %.17g %.17g\n" % c for c in cases)
        out = subprocess.run([str(exe)], input=inp, capture_output=True, text=True, check=True).stdout
        return [int(v) for v in out.split()]
    return run


def _bf16_lossless(v):
    """True when every value of v is a BF16 value (8 significant bits)."""
    u = np.asarray(v, np.float32).view(np.uint32)
    return bool((u & 0xFFFF == 0).all())


def test_bf16_values():
    assert _bf16_lossless([0, 1, 255, 256, 258, 700, 512, 65280])
    assert not _bf16_lossless([257]) and not _bf16_lossless([701]) and not _bf16_lossless([699])


L2, EUC, ANG = 0, 1, 2


def test_exact_cases(predicate):
    cases = [
        # the 1M x 128 integer mixture: values in [0, 255]
        ((128, L2, 1, 1, 255, 0, 255), 1),
        ((128, EUC, 1, 1, 255, 0, 255), 1),
        # D = 16 rows up to 700: the sums fit, but 701 is not a BF16 value
        ((16, L2, 1, 0, 700, 0, 700), 0),
        # the same magnitudes when every value is a BF16 value (multiples of 4)
        ((16, L2, 1, 1, 700, 0, 700), 1),
        # 2 D max^2 = 2^24 + 4864 at D = 16, max = 725: a norm sum can round
        ((16, L2, 1, 1, 725, 0, 725), 0),
        # both signs: |a - o| reaches 2 max, D (2 max)^2 > 2^24 while 2 D max^2 <= 2^24
        ((128, L2, 1, 1, 200, -200, 200), 0),
        ((128, EUC, 1, 1, 200, -200, 200), 0),
        # the angular form never forms a difference: only the dot must be exact
        ((128, ANG, 1, 1, 200, -200, 200), 1),
        # non-integer rows are never exact
        ((16, L2, 0, 1, 0.5, 0, 0.5), 0),
        # negative-only rows with a small range
        ((128, L2, 1, 1, 255, -255, 0), 1),
    ]
    assert predicate([c for c, _ in cases]) == [e for _, e in cases]


def _tile_and_canonical_f32(a, o, metric):
    """FP32 value of the tile formula and of the canonical sequential form,
    each operation rounded to FP32 (numpy float32 scalars)."""
    f = np.float32
    na = f(0); no = f(0); g = f(0); ssd = f(0)
    for x, y in zip(a.astype(np.float32), o.astype(np.float32)):
        na = f(na + x * x); no = f(no + y * y); g = f(g + x * y)
        d = f(x - y); ssd = f(ssd + d * d)
    if metric == L2:
        return f(f(na + no) - f(2) * g), ssd
    return f(f(f(na * f(0.5)) + f(no * f(0.5))) - g), f(ssd * f(0.5))


def test_predicate_is_sufficient_on_extreme_rows(predicate):
    # where the predicate says "exact", the extreme rows (all max, and max
    # against min) give the exact integer result through both FP32 forms
    rs = np.random.RandomState(0)
    cases, rows = [], []
    while len(cases) < 200:
        D = int(rs.randint(1, 257))
        mx = int(rs.randint(1, 1500))
        lo = -int(rs.randint(0, mx + 1)) if rs.rand() < 0.5 else 0
        metric = int(rs.randint(0, 2))
        cases.append((D, metric, 1, 1, float(mx), float(lo), float(mx)))
        rows.append((D, metric, mx, lo))
    got = predicate(cases)
    assert 0 < sum(got) < len(got)
    for ok, (D, metric, mx, lo) in zip(got, rows):
        if not ok:
            continue
        for a, o in ((np.full(D, mx), np.full(D, mx)), (np.full(D, mx), np.full(D, lo)),
                     (np.full(D, lo), np.full(D, lo))):
            exact = float(((a.astype(np.int64) - o) ** 2).sum()) * (1.0 if metric == L2 else 0.5)
            tile, canon = _tile_and_canonical_f32(a, o, metric)
            assert float(tile) == exact and float(canon) == exact, (D, metric, mx, lo)


PROGRAM_I8 = r'''
#include <cstdio>
#include "kg_exact.hpp"
int main() {
    unsigned D; int metric, u8;
    while (std::scanf("%u %d %d", &D, &metric, &u8) == 3)
        std::printf("%d\n", kga::i8_tiles_exact(D, metric, u8 != 0) ? 1 : 0);
    return 0;
}
'''


@pytest.fixture(scope="module")
def predicate_i8(tmp_path_factory):
    d = tmp_path_factory.mktemp("exact_i8")
    src, exe = d / "p.cpp", d / "p"
    src.write_text(PROGRAM_I8)
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + CSRC, str(src), "-o", str(exe)], check=True)

    def run(cases):
        inp = "".join("%d %d %d\n" % c for c in cases)
        out = subprocess.run([str(exe)], input=inp, capture_output=True, text=True, check=True).stdout
        return [int(v) for v in out.split()]
    return run


def test_i8_exact_cases(predicate_i8):
    # 8-bit rows (integers in [0, 255]): exact while 255^2 D <= 2^24
    cases = [
        ((128, L2, 1), 1), ((128, EUC, 1), 1),
        ((200, L2, 1), 1),    # 2 D 255^2 > 2^24: the BF16 form is not exact here, the 8-bit one is
        ((256, L2, 1), 1),    # 255^2 256 = 16,646,400
        ((258, EUC, 1), 1),   # 255^2 258 = 16,776,450
        ((259, L2, 1), 0),    # 255^2 259 = 16,841,475 > 2^24
        ((128, ANG, 1), 0),   # the angular distance is not shift-invariant
        ((128, L2, 0), 0),    # no exact 8-bit copy
    ]
    assert predicate_i8([c for c, _ in cases]) == [e for _, e in cases]


def _i8_tile_and_canonical_f32(a, o, metric):
    """The 8-bit form's FP32 tile value (exact int32 norms and dot of the rows
    shifted by -128, then (n_a + n_o) - 2 G, or the halved form) and a
    sequential FP32 sum of squared differences (every partial sum is an
    integer FP32 holds here, so any summation order gives this value)."""
    f = np.float32
    sa, so = a.astype(np.int64) - 128, o.astype(np.int64) - 128
    na, no, g = int((sa * sa).sum()), int((so * so).sum()), int((sa * so).sum())
    ssd = f(0)
    for x, y in zip(a.astype(np.float32), o.astype(np.float32)):
        d = f(x - y)
        ssd = f(ssd + d * d)
    if metric == L2:
        return f(f(f(na) + f(no)) - f(2) * f(g)), ssd
    return f(f(f(na) * f(0.5) + f(no) * f(0.5)) - f(g)), f(ssd * f(0.5))


def test_i8_predicate_is_sufficient(predicate_i8):
    # where the 8-bit predicate says "exact", extreme rows (all 255 against all
    # 0, alternating) and random 8-bit rows give the exact integer through the
    # tile formula and the canonical order
    rs = np.random.RandomState(1)
    Ds = [1, 3, 37, 64, 100, 128, 200, 255, 256, 258]
    got = predicate_i8([(D, m, 1) for D in Ds for m in (L2, EUC)])
    assert all(got)
    for D in Ds:
        for metric in (L2, EUC):
            alt = np.where(np.arange(D) % 2 == 0, 255, 0)
            pairs = [(np.full(D, 255), np.zeros(D, np.int64)), (alt, 255 - alt), (np.zeros(D, np.int64),
                                                                                  np.zeros(D, np.int64))]
            pairs += [(rs.randint(0, 256, D), rs.randint(0, 256, D)) for _ in range(5)]
            for a, o in pairs:
                exact = float(((a.astype(np.int64) - o) ** 2).sum()) * (1.0 if metric == L2 else 0.5)
                tile, canon = _i8_tile_and_canonical_f32(a, o, metric)
                assert float(tile) == exact and float(canon) == exact, (D, metric)
