"""Plans driven by the reference's spherical grid / mode files
(sht_sph_files_*): the reader parses sets written by the reference's own
writers with its own ordering routines (the committed fixture, and fresh sets
from oracle/_ref where it is built), the one-rank check accepts them, and
every kind of difference from the one-rank layout is reported with the file
and the entry.  CPU only: no plan is created here."""
import os
import shutil

import numpy as np
import pytest

from kemorin_mhd_amd import ShellConfig, ShtError, SphFiles, rlm_l_m
from tests.golden.make_sph_files import SHAPE, SPACES, paths as fixture_paths

try:
    from oracle import ref
    HAVE_REF = ref.available()
except Exception:  # pragma: no cover - the reference build is optional
    HAVE_REF = False


def _fft_slot(m, nth):
    return 1 if m == 0 else 2 if m == nth else 2 * m + 1 if m > 0 else -2 * m + 2


def _expect_tables(c: ShellConfig):
    """The one-rank tables of CONST_SPH_GRID/set_sph_tranform_ordering.f90,
    enumerated independently of the library."""
    ls, ms = rlm_l_m(c.L)
    modes = np.stack([ls * (ls + 1) + ms, ls, ms], axis=1).ravel()
    slots = np.array([[_fft_slot(m, c.nth), m] for m in range(-c.L, c.L + 1)]).ravel()
    inv = {_fft_slot(m, c.nth): m for m in range(-c.nth + 1, c.nth + 1)}
    phi = np.array([[i, inv[i]] for i in range(1, c.nphi + 1)]).ravel()
    return modes, slots, phi


def _check_set(f: SphFiles, c: ShellConfig, radius):
    cfg, r, nnod_rj = f.shape()
    assert (cfg.L, cfg.nr, cfg.nth, cfg.nphi) == (c.L, c.nr, c.nth, c.nphi)
    assert nnod_rj == c.jmax * c.nr
    # radii as written: 1pE25.15 keeps 16 significant digits
    assert np.array_equal(r, [float(f"{x:.15e}") for x in radius])
    f.check()
    modes, slots, phi = _expect_tables(c)
    assert np.array_equal(f.table("rlm_table2"), modes)
    assert np.array_equal(f.table("rj_table2"), modes)
    assert np.array_equal(f.table("rtm_table3"), slots)
    assert np.array_equal(f.table("rtp_table3"), phi)
    assert np.array_equal(f.table("rtp_table2"), np.arange(1, c.nth + 1))
    n = c.jmax * c.nr
    assert np.array_equal(f.table("rj_import"), np.arange(1, n + 1))
    assert np.array_equal(f.table("rj_import_rev"), np.arange(1, n + 1))
    # rtp import: rtm node (slot s, k, l) -> rtp node of (k, l, FFT slot of s)
    s, k, l = np.meshgrid(np.arange(c.nslot), np.arange(c.nr), np.arange(c.nth), indexing="ij")
    fs = slots.reshape(-1, 2)[:, 0][s]
    want = 1 + k + c.nr * l + c.nr * c.nth * (fs - 1)
    assert np.array_equal(f.table("rtp_import"), want.ravel())


def test_fixture_reads_and_checks():
    f = SphFiles(*fixture_paths())
    try:
        _check_set(f, SHAPE, SHAPE.radius())
    finally:
        f.close()


@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref (the reference's writers) not built")
@pytest.mark.parametrize("shape", [(3, 2, 4, 8), (15, 5, 24, 48), (12, 3, 13, 26)])
def test_sets_written_by_the_reference(tmp_path, shape):
    c = ShellConfig(*shape)
    ps = [str(tmp_path / f"in.{s}") for s in SPACES]
    ref.write_sph_files(c.L, c.nr, c.nth, c.nphi, c.radius(), ps)
    f = SphFiles(*ps)
    try:
        _check_set(f, c, c.radius())
    finally:
        f.close()


def _copy_fixture(tmp_path):
    out = []
    for p in fixture_paths():
        q = tmp_path / os.path.basename(p)
        shutil.copy(p, q)
        out.append(str(q))
    return out


def _edit(path, old, new, count=1):
    s = open(path).read()
    assert old in s
    open(path, "w").write(s.replace(old, new, count))


def _lines(path):
    return open(path).read().split("\n")


def _check_fails(ps, *words):
    f = SphFiles(*ps)
    try:
        with pytest.raises(ShtError) as e:
            f.check()
        for w in words:
            assert w in str(e.value), str(e.value)
    finally:
        f.close()


def test_rejects_mode_table_difference(tmp_path):
    ps = _copy_fixture(tmp_path)
    rlm = ps[2]
    lines = _lines(rlm)
    # the first mode entry (l = L, m = -L) becomes (l = L, m = -L + 1)
    i = next(i for i, x in enumerate(lines) if x.split() == ["25", "5", "-5"])
    lines[i] = lines[i].replace("-5", "-4")
    open(rlm, "w").write("\n".join(lines))
    _check_fails(ps, "in.rlm", "mode 1")


def _import_start(lines):
    """Line of the first import item: the run 1, 2, 3 of one-value lines
    after the node list."""
    return max(i for i in range(len(lines) - 2)
               if [lines[i + d].split() for d in range(3)] == [["1"], ["2"], ["3"]])


def test_rejects_permuted_rlm_import(tmp_path):
    ps = _copy_fixture(tmp_path)
    lines = _lines(ps[2])
    a = _import_start(lines)
    lines[a], lines[a + 1] = lines[a + 1], lines[a]
    open(ps[2], "w").write("\n".join(lines))
    _check_fails(ps, "in.rlm", "import item 1")


def test_rejects_unpaired_rj_import(tmp_path):
    ps = _copy_fixture(tmp_path)
    lines = _lines(ps[3])
    a = _import_start(lines)
    lines[a], lines[a + 1] = lines[a + 1], lines[a]
    open(ps[3], "w").write("\n".join(lines))
    _check_fails(ps, "in.rj", "radius and mode")


def test_rejects_folded_phi_table(tmp_path):
    ps = _copy_fixture(tmp_path)
    lines = _lines(ps[0])
    # phi entry 3 (slot 3 = m 1) claims m = 2, as an m_folding = 2 set would
    i = next(i for i, x in enumerate(lines) if x.split() == ["3", "1"])
    lines[i] = lines[i].replace(" 1", " 2")
    open(ps[0], "w").write("\n".join(lines))
    _check_fails(ps, "in.rtp", "phi entry 3")


def test_rejects_multi_rank_set(tmp_path):
    ps = _copy_fixture(tmp_path)
    lines = _lines(ps[1])
    # subdomain index of the m direction = 1 (a second rank's share)
    i = next(i for i, x in enumerate(lines) if x.split() == ["0", "0", "0"])
    lines[i] = "               0               0               1"
    open(ps[1], "w").write("\n".join(lines))
    _check_fails(ps, "in.rtm", "not a one-rank set")


def test_truncated_file_names_file_and_line(tmp_path):
    ps = _copy_fixture(tmp_path)
    lines = _lines(ps[3])
    open(ps[3], "w").write("\n".join(lines[:40]))
    with pytest.raises(ShtError, match=r"in\.rj:\d+:"):
        SphFiles(*ps)


def test_missing_file_and_unknown_table(tmp_path):
    ps = _copy_fixture(tmp_path)
    with pytest.raises(ShtError, match="cannot open"):
        SphFiles(ps[0], ps[1], ps[2], str(tmp_path / "absent"))
    f = SphFiles(*ps)
    try:
        with pytest.raises(ShtError):
            f.table("rj_nonsense")
    finally:
        f.close()
