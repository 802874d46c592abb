"""Checkpoint and resume formats: raw FP64 gauge files (gauge_ext.h:225-274), the RNG
text state of ParallelRngExt (rng.h:28-52, 131-168) and the resume scan (hmc.cu:423-447)."""
import numpy as np
import pytest

import qed3_amd as q
from qed3_amd import _lib

M64 = (1 << 64) - 1


def mt64_seed(seed):
    """mt19937_64 seeding recursion of the C++ standard ([rand.eng.mers]): 312 words."""
    x = [seed & M64]
    for i in range(1, 312):
        x.append((6364136223846793005 * (x[-1] ^ (x[-1] >> 62)) + i) & M64)
    return x


def mt64_outputs(seed, k):
    """First k outputs of mt19937_64 (twist and tempering of the standard)."""
    x, out, i = mt64_seed(seed), [], 312
    for _ in range(k):
        if i >= 312:
            for j in range(312):
                y = (x[j] & 0xFFFFFFFF80000000) | (x[(j + 1) % 312] & 0x7FFFFFFF)
                v = x[(j + 156) % 312] ^ (y >> 1)
                x[j] = v ^ (0xB5026F5AA96619E9 if y & 1 else 0)
            i = 0
        y = x[i]
        i += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        out.append(y & M64)
    return out


def as_int_seed(u):
    """SingleRng::reseed(const int): the master draw narrowed to int, then widened back."""
    v = u & 0xFFFFFFFF
    v = v - (1 << 32) if v >= (1 << 31) else v
    return v & M64


def test_python_mt64_known_answer():
    assert mt64_outputs(5489, 10000)[-1] == 9981545732273789042


def test_gauge_file_roundtrip_and_layout(tmp_path):
    Lt, E, V = 3, 30, 12
    rng = np.random.default_rng(1)
    sp, tp = rng.normal(size=(Lt, E)), rng.normal(size=(Lt, V))
    f = tmp_path / "ckpoint_lat.10"
    q.gauge_write(f, sp, tp)
    raw = np.fromfile(f, dtype=np.float64)
    assert raw.size == Lt * (E + V)
    assert np.array_equal(raw, np.concatenate([sp.ravel(), tp.ravel()]))
    sp2, tp2 = q.gauge_read(f, Lt, E, V)
    assert np.array_equal(sp2, sp) and np.array_equal(tp2, tp)
    with pytest.raises(_lib.Qed3Error):
        q.gauge_read(f, Lt, E, V + 1)  # too few doubles
    with pytest.raises(_lib.Qed3Error):
        q.gauge_read(f, Lt, E - 1, V)  # too many


def test_rng_file_layout(tmp_path):
    Lt, E, V, seed = 2, 5, 3, 7
    r = q.Rng(Lt, E, V, seed)
    f = tmp_path / "ckpoint_rng.10"
    r.save(f)
    lines = f.read_text().splitlines()
    assert len(lines) == 4 * (1 + Lt * E + Lt * V)
    master = mt64_outputs(seed, Lt * (E + V))
    # link stream (t=0, l=0) is the second block: an untouched engine seeded with the
    # first master draw, written as its 312 state words and the position 312
    words = [int(w) for w in lines[4].split()]
    assert words == mt64_seed(as_int_seed(master[0])) + [312]
    # the first site stream follows all Lt*E link streams
    words = [int(w) for w in lines[4 * (1 + Lt * E)].split()]
    assert words == mt64_seed(as_int_seed(master[Lt * E])) + [312]
    # distribution lines parse as numbers (normal, uniform [0,1), int {1,2})
    for k in (1, 2, 3):
        [float(t) for t in lines[k].split()]
    assert [float(t) for t in lines[2].split()] == [0.0, 1.0]
    assert [int(t) for t in lines[3].split()] == [1, 2]


def test_rng_resume_matches_the_reference_reader(tmp_path):
    """Loading restores the engines and resets the distributions: after an even number of
    gaussian draws the resumed stream continues exactly; after an odd number the pending
    second value of the polar method is dropped, as in a resumed reference run."""
    Lt, E, V = 2, 4, 3
    for ndraw, skip in ((2, 0), (1, 1)):
        a = q.Rng(Lt, E, V, 11)
        for _ in range(ndraw):
            a.gaussian(q.Rng.SITE, 1, 2)
            a.gaussian(q.Rng.LINK, 0, 3)
            a.gaussian(q.Rng.MASTER)
        f = tmp_path / f"r{ndraw}"
        a.save(f)
        cont = [[a.gaussian(w, s, i) for _ in range(4)]
                for w, s, i in ((q.Rng.SITE, 1, 2), (q.Rng.LINK, 0, 3), (q.Rng.MASTER, 0, 0))]
        b = q.Rng(Lt, E, V, 99)  # other seed: every stream comes from the file
        b.load(f)
        res = [[b.gaussian(w, s, i) for _ in range(3)]
               for w, s, i in ((q.Rng.SITE, 1, 2), (q.Rng.LINK, 0, 3), (q.Rng.MASTER, 0, 0))]
        for c, r in zip(cont, res):
            assert r == c[skip:skip + 3]
        # untouched streams continue exactly either way
        fresh = q.Rng(Lt, E, V, 11)
        assert b.gaussian(q.Rng.SITE, 0, 0) == fresh.gaussian(q.Rng.SITE, 0, 0)


def test_rng_load_rejects_a_truncated_file(tmp_path):
    r = q.Rng(2, 4, 3, 1)
    f = tmp_path / "r"
    r.save(f)
    f.write_text("\n".join(f.read_text().splitlines()[:-5]) + "\n")
    with pytest.raises(_lib.Qed3Error):
        q.Rng(2, 4, 3, 1).load(f)


def test_resume_scan(tmp_path):
    assert q.resume_point(tmp_path) == 0
    for k in (10, 20, 40):
        for p in q.checkpoint_paths(tmp_path, k):
            open(p, "w").close()
    assert q.resume_point(tmp_path) == 20  # stops at the first gap (30)
    import os
    os.remove(q.checkpoint_paths(tmp_path, 20)[1])  # rng file of 20 missing
    assert q.resume_point(tmp_path) == 10


@pytest.mark.gpu
def test_hmc_resumes_from_a_checkpoint(gpu, tmp_path):
    """A fresh HMC handle (other seed) that loads the RNG file and the gauge file runs the
    same trajectory as the handle that wrote them and reloaded its own streams."""
    from pairs import links
    lattice, n, Lt = 0, 2, 4
    D = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=0.1, at=0.2)
    sp, tp = links(lattice, n, Lt, seed=3)
    sp, tp = np.ascontiguousarray(sp), np.ascontiguousarray(tp)
    a = q.HMC(D, gsq=1 / 1.4, npf=1, nsteps=4, nsteps_inner=2, seed=5)
    a.trajectory(sp, tp, no_reject=True)
    lat, rng = q.checkpoint_paths(tmp_path, 10)
    q.gauge_write(lat, sp, tp)
    a.rng_save(rng)
    a.rng_load(rng)
    ra = a.trajectory(sp, tp, no_reject=True)
    b = q.HMC(D, gsq=1 / 1.4, npf=1, nsteps=4, nsteps_inner=2, seed=77)
    b.rng_load(rng)
    sp2, tp2 = q.gauge_read(lat, Lt, D.E, D.V)
    rb = b.trajectory(sp2, tp2, no_reject=True)
    assert np.array_equal(sp2, sp) and np.array_equal(tp2, tp)
    assert rb["dH"] == ra["dH"]
