"""The roofline denominators of bench.py: algorithmic bytes per launch of the D stencil and
of the CG kernels, checked against figures counted by hand for dual n=128, Lt=64
(V = 20 n^2 = 327,680 sites, E = 30 n^2 = 491,520 links) and simplicial n=128, Lt=64
(V = 10 n^2 + 2, E = 30 n^2)."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402

V_DUAL, E_DUAL, LT = 327_680, 491_520, 64
PTS = V_DUAL * LT  # 20,971,520 points


def test_bare_d_bytes():
    # one field read + one written (2 complex per point), one real per link (E + V per slice)
    for real_bytes in (2, 4, 8):
        field = PTS * 2 * 2 * real_bytes
        link = (E_DUAL + V_DUAL) * LT * real_bytes
        assert bench.algorithmic_bytes_D(V_DUAL, E_DUAL, LT, real_bytes) == 2 * field + link
    assert bench.algorithmic_bytes_D(V_DUAL, E_DUAL, LT, 2) == 440_401_920  # binary16 D
    assert bench.algorithmic_bytes_D(V_DUAL, E_DUAL, LT, 8) == 1_761_607_680  # FP64 D


def test_half_cg_kernel_bytes():
    # binary16 fields: 8 B per point; FP32 x: 16 B per point; the temporal gauge leaves the
    # E Lt spatial 16-bit angles and one wrap angle per site
    b = bench.cg_kernel_bytes(V_DUAL, E_DUAL, LT, "half")
    vec, x, link = PTS * 8, PTS * 16, (E_DUAL * LT + V_DUAL) * 2
    assert link == 63_569_920
    assert b["K1"] == 4 * vec + link == 734_658_560  # r, p_old read; p, w written
    assert b["K2"] == 3 * vec + link == 566_886_400  # w, r read; r written
    assert b["K2x"] == 7 * vec + 2 * x + link == 1_909_063_680  # + 4 directions, x read/written


def test_fp_cg_kernel_bytes():
    for solver, s in (("mixed", 4), ("fp64", 8)):
        b = bench.cg_kernel_bytes(V_DUAL, E_DUAL, LT, solver)
        vec, link = PTS * 2 * 2 * s, (E_DUAL + V_DUAL) * LT * s
        assert b["K1"] == 4 * vec + link
        assert b["K2"] == 6 * vec + link
    assert bench.cg_kernel_bytes(V_DUAL, E_DUAL, LT, "mixed")["K1"] == 1_551_892_480


def test_simplicial_counts():
    n = 128
    V, E = 10 * n * n + 2, 30 * n * n
    assert bench.algorithmic_bytes_D(V, E, LT, 4) == 2 * (V * LT * 16) + (E + V) * LT * 4


def test_preconditioner_layer_bytes():
    # a transport layer reads the injection and its input field and writes its output
    # (binary16, 8 B per point each), reads the spatial link phases as 10-bit angles (10 bits
    # per link and slice: 491,520 x 64 x 10 / 8 = 39,321,600 B) and one wrap phase per site as
    # a binary16 pair (4 B), and the per-site tables once: 3 slots x (2 spin rows x 8 B hop
    # coefficients + 2 B local index) + 1/c + kappa_t
    b = bench.layer_bytes(V_DUAL, E_DUAL, LT, 3, False)
    assert b == 503_316_480 + 39_321_600 + 1_310_720 + 20_316_160 == 564_264_960
    assert PTS == 20_971_520 and V_DUAL * (3 * 18 + 8) == 20_316_160
    # a first layer: the injection read and the output written, 1/c and kappa_t
    assert bench.layer_bytes(V_DUAL, E_DUAL, LT, 3, True) == 2 * PTS * 8 + V_DUAL * 8 == 338_165_760
