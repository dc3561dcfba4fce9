"""The strip schedule's launch shapes (mcrg::strip_shape, csrc/strip.hpp) on
the CPU, through tests/cpp/strip_shape_probe, for every lattice size, level
count and mode the runner can launch: each workspace region holds the largest
index its kernel uses, regions do not overlap, every thread has at most the
words per exchange the kernels assume, the slot counts keep the race-freedom
bounds the kernels rely on, and the allocation-free entry point's static
array holds every shape it runs.  No GPU."""
import math
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROBE = os.path.join(ROOT, "mcrg_amd", "_build", "strip_shape_probe")
LDS_BYTES = 160 * 1024


def _run(lines):
    if not os.path.exists(PROBE):
        pytest.skip("mcrg_amd/_build/strip_shape_probe not built (make -C mcrg_amd)")
    out = subprocess.run([PROBE], input="".join(lines), capture_output=True, text=True, check=True).stdout
    consts, shapes = {}, []
    for line in out.splitlines():
        kind, *fields = line.split()
        d = {k: int(v) for k, v in (f.split("=") for f in fields)}
        if kind == "const":
            consts = d
        else:
            shapes.append(d)
    return consts, shapes


def _cases():
    # the runner: n_wg = min(256, L/2) strips (256 compute units), tails 4
    # (fewer only when they do not fit), halo rows and measuring workgroups
    # allowed unless switched off
    lines = []
    for lg in range(6, 15):
        L = 1 << lg
        n_wg = min(256, L // 2)
        for measure in (0, 1):
            if measure and L > 8192:
                continue
            for nl in range(0, lg - 1):
                for halo_ok in (0, 1):
                    for split_ok in (0, 1):
                        for n_tail in (4, 2, 1):
                            lines.append(f"{L} {nl} {n_wg} {n_tail} {measure} {halo_ok} {split_ok}\n")
    return lines


@pytest.fixture(scope="module")
def shapes():
    return _run(_cases())


def _accepted(s):
    # the runner's own checks (StripRunner::plan): one neighbour word of each
    # kind per thread, and the strip's LDS within a compute unit's
    if (4 if s["halo"] else 2) * s["wpr"] > s["threads"] or s["first_words"] > s["threads"]:
        return False
    return s["lds_words"] * 4 <= LDS_BYTES


def test_every_shape_the_runner_launches_fits_its_checks(shapes):
    _, sh = shapes
    # every lattice size the strip schedule takes is launched in some mode
    for L in (64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384):
        assert any(_accepted(s) and s["L"] == L for s in sh), L


def test_workspace_regions_hold_the_kernels_indices(shapes):
    c, sh = shapes
    for s in sh:
        n, wpr = s["n_wg"], s["wpr"]
        assert s["o_err"] == 0 and s["o_consumed"] >= 1
        # per-slot consumed counters, kStripFlagStride words apart
        assert s["o_xhalo"] - s["o_consumed"] >= c["kStripRing"] * c["kStripFlagStride"]
        # halo rows: (data, seq) = 2 words each; [parity][strip][first two,
        # last two rows] with halo, [colour][strip][first, last row] without
        halo_u64 = 2 * n * (4 if s["halo"] else 2) * wpr
        assert s["o_xfirst"] - s["o_xhalo"] >= 2 * halo_u64
        assert s["o_ring"] - s["o_xfirst"] >= 2 * s["first_slots"] * n * s["first_words"]
        if s["measure"]:
            assert s["o_snapc"] - s["o_ring"] >= 2 * c["kStripRing"] * n * s["per_wg"]
        if s["split"]:
            assert s["o_snap"] - s["o_snapc"] >= n * c["kStripFlagStride"]
            assert s["ws_words"] - s["o_snap"] >= 2 * c["kStripSnapSlots"] * n * s["snap_words"]
        offs = [s[k] for k in ("o_err", "o_consumed", "o_xhalo", "o_xfirst", "o_ring", "o_snapc", "o_snap",
                               "ws_words")]
        assert offs == sorted(offs)
        # 64-bit words start at even 32-bit offsets
        for k in ("o_xhalo", "o_xfirst", "o_ring", "o_snap"):
            assert s[k] % 2 == 0, (k, s)


def test_levels_and_rows(shapes):
    _, sh = shapes
    for s in sh:
        R, L = s["R"], s["L"]
        assert R == L // s["n_wg"] and R >= 2 and s["wpr"] == L // 64
        if not s["measure"]:
            assert s["lam"] == 0 and s["first_words"] == 0 and s["n_tail"] == 0 and not s["split"]
            continue
        assert s["lam"] == min(int(math.log2(R)), s["nl"])
        # level l's row: (L >> l) / 32 words, packed one after another
        off = 0
        for l in range(1, s["lam"] + 1):
            assert s[f"row_off{l}"] == off
            off += (L >> l) // 32
        assert s["first_words"] == off
        assert s["cnt_words"] == 4 * (s["lam"] + 1)
        assert s["lam_words"] == ((L >> s["lam"]) // 32 if s["nl"] > s["lam"] else 0)
        assert s["per_wg"] == s["cnt_words"] + s["lam_words"]


def test_words_per_thread(shapes):
    _, sh = shapes
    for s in filter(_accepted, sh):
        R, wpr, nt = s["R"], s["wpr"], s["threads"]
        assert nt in (64, 256, 512, 1024)
        if s["halo"]:
            assert (R + 2) * wpr <= 512 and not s["overlap"]
            # colour 0: one word per thread; colour 1: one, or with measuring
            # workgroups at most two (words j and j + threads)
            assert R * wpr <= nt
            assert (R + 2) * wpr <= (2 if s["split"] else 1) * nt
        if s["split"]:
            assert s["halo"] and s["measure"] and nt in (64, 256)
            assert s["n_meas"] == s["n_wg"] and s["snap_words"] == 2 * (R + 2) * wpr
        else:
            assert s["n_meas"] == 0 and s["snap_words"] == 0
        if s["overlap"]:
            assert s["measure"] and R * wpr <= nt


def test_slot_counts_keep_the_race_bounds(shapes):
    c, sh = shapes
    for s in sh:
        if s["split"]:
            # a measuring workgroup's first rows are read within
            # kStripSnapSlots + 2 samples of being written
            assert s["first_slots"] >= c["kStripSnapSlots"] + 3
        elif s["halo"]:
            # read a sweep after they are sent; the sender may send two more
            assert s["first_slots"] >= 3
        else:
            assert s["first_slots"] == 1


def test_static_array_holds_every_shape_without_measuring_workgroups(shapes):
    c, sh = shapes
    worst = max(s["ws_words"] for s in filter(_accepted, sh) if not s["split"])
    assert worst <= c["kStaticScratchWords"]
    assert worst * 4 <= 2.2 * 2**20  # "at most 2.1 MiB"


def test_halo_and_split_are_chosen_where_they_fit(shapes):
    _, sh = shapes
    for s in filter(_accepted, sh):
        hw = (s["R"] + 2) * s["wpr"]
        key = (s["L"], s["nl"], s["measure"])
        if hw > 512:
            assert not s["halo"], key
    # the C2 shape: halo rows and measuring workgroups on 256-thread workgroups
    c2 = [s for s in sh if s["L"] == 2048 and s["nl"] == 5 and s["measure"] and s["n_tail"] == 4]
    assert any(s["split"] and s["threads"] == 256 and s["lam"] == 3 for s in c2)
