"""The autotune table file (no GPU needed): entries round-trip through save and load,
every tile width the fused kernels have is accepted (256 included), and a file with a
bad entry is refused whole, leaving the table as it was."""
import pytest

import flashgemm_amd as fg

HEADER = "flashgemm-autotune v2 gfx950\n"
# kind num_gemm N-bucket M-buckets K-buckets segments (end tile)... splits ms
GOOD = [
    "2 3 14 9 10 9 9 9 10 1 2 64 0 0.033",           # INT8 3-chain fused at 64 columns
    "1 2 13 8 6 8 8 1 1 256 0 0.010",                 # BF16 2-chain fused at 256 columns
    "0 4 14 10 12 10 12 10 10 12 10 3 1 32 2 0 3 0 0 1.5",  # FP32: GEMMs 0-1 fused, 2 and 3 unfused
    "3 2 11 7 7 7 7 2 0 0 1 0 1 0.2",                 # split-bf16 x3: both GEMMs unfused, split-K off
]


def lib_or_skip():
    try:
        return fg.lib()
    except RuntimeError as e:
        pytest.skip(str(e))


def write(tmp_path, name, lines, header=HEADER):
    p = tmp_path / name
    p.write_text(header + "".join(l + "\n" for l in lines))
    return str(p)


def test_roundtrip(tmp_path):
    lib_or_skip()
    fg.autotune_clear()
    try:
        fg.autotune_load(write(tmp_path, "good.txt", GOOD))
        assert fg.autotune_size() == len(GOOD)
        out = str(tmp_path / "saved.txt")
        fg.autotune_save(out)
        fg.autotune_clear()
        assert fg.autotune_size() == 0
        fg.autotune_load(out)
        assert fg.autotune_size() == len(GOOD)
        out2 = str(tmp_path / "saved2.txt")
        fg.autotune_save(out2)
        assert open(out).read() == open(out2).read()
        assert open(out).read().startswith(HEADER)
    finally:
        fg.autotune_clear()


@pytest.mark.parametrize("bad", [
    "2 2 10 6 6 6 6 1 1 96 0 0.1",        # INT8 has no 96-column tile
    "0 2 10 6 6 6 6 1 1 48 0 0.1",        # not a tile width at all
    "0 3 10 6 6 6 6 6 6 2 1 0 2 32 0 0.1",  # an unfused segment of two GEMMs
    "0 2 10 6 6 6 6 2 1 0 0 0 0 0.1",     # segment ends out of order
    "0 2 10 6 6 6 6 1 0 64 0 0.1",        # segments that stop before the last GEMM
    "4 1 10 6 6 1 0 0 0 0.1",             # no such kind
    "0 9 10 6 6 6 6 6 6 6 6 6 6 6 6 6 6 6 6 6 6 1 8 32 0 0.1",  # nine GEMMs
    "0 2 10 6 6 6 6 1 1 64 -1 0.1",       # negative split-K
    "0 2 10 6 6",                          # truncated
])
def test_bad_entry_refused_table_unchanged(tmp_path, bad):
    lib_or_skip()
    fg.autotune_clear()
    try:
        fg.autotune_load(write(tmp_path, "good.txt", GOOD[:1]))
        with pytest.raises(fg.FlashgemmError) as e:
            fg.autotune_load(write(tmp_path, "bad.txt", GOOD[1:] + [bad]))
        assert e.value.status == fg.ERR_INVALID
        assert fg.autotune_size() == 1
    finally:
        fg.autotune_clear()


def test_other_versions_refused(tmp_path):
    lib_or_skip()
    fg.autotune_clear()
    for header in ("flashgemm-autotune v1 gfx950\n", "flashgemm-autotune v2 gfx942\n", "something else\n"):
        with pytest.raises(fg.FlashgemmError):
            fg.autotune_load(write(tmp_path, "old.txt", GOOD, header=header))
    assert fg.autotune_size() == 0
